"""ctypes binding of libcabanamd_amd.so (the C ABI in include/cabanamd_amd.h).

The library must be built (python -m cabanamd_amd.build); loading fails
loudly otherwise -- there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcabanamd_amd.so")

DP = C.POINTER(C.c_double)
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int)
LP = C.POINTER(C.c_int64)
VP = C.c_void_p

CMD_OK = 0
CMD_ERR_CAPACITY = 5
CMD_ERR_NO_DEVICE = 7
PREC_FP64 = 0
PREC_MIXED = 1
MAX_TYPES = 8
MAX_REGIONS = 16


class CmdError(RuntimeError):
    """Raised for any non-zero cmd_status (the reference's log_err throws
    std::runtime_error, output.h:37-54)."""

    def __init__(self, status, msg):
        super().__init__(f"cmd_status {status}: {msg}")
        self.status = status


class Domain(C.Structure):
    _fields_ = [("low", C.c_double * 3), ("high", C.c_double * 3),
                ("L", C.c_double * 3), ("cell", C.c_double * 3),
                ("local_lo", C.c_double * 3), ("local_hi", C.c_double * 3),
                ("local_mesh", C.c_double * 3),
                ("ghost_lo", C.c_double * 3), ("ghost_hi", C.c_double * 3),
                ("halo_width", C.c_int)]


class Params(C.Structure):
    _fields_ = [("units", C.c_int), ("lattice_style", C.c_int),
                ("lattice_constant", C.c_double),
                ("lattice_offset", C.c_double * 3),
                ("region_lo", C.c_double * 3), ("region_hi", C.c_double * 3),
                ("ntypes", C.c_int), ("mass", C.c_double * MAX_TYPES),
                ("temp_target", C.c_double), ("temp_seed", C.c_int),
                ("vel_gaussian", C.c_int),
                ("force_cutoff", C.c_double), ("skin", C.c_double),
                ("every", C.c_int), ("max_neigh_guess", C.c_int),
                ("comm_ghost_cutoff", C.c_double), ("half", C.c_int),
                ("nsteps", C.c_int), ("thermo_rate", C.c_int),
                ("dt", C.c_double), ("mvv2e", C.c_double),
                ("boltz", C.c_double), ("ncoeff", C.c_int),
                ("coeff_i", C.c_int * 16), ("coeff_j", C.c_int * 16),
                ("coeff_eps", C.c_double * 16),
                ("coeff_sigma", C.c_double * 16),
                ("coeff_cut", C.c_double * 16),
                ("force_type", C.c_int), ("nnp_dir", C.c_char * 256),
                ("atom_style_charge", C.c_int), ("data_in", C.c_char * 256),
                ("data_out", C.c_char * 256),
                ("nregions", C.c_int),
                ("region_name", (C.c_char * 32) * MAX_REGIONS),
                ("region_block", (C.c_double * 6) * MAX_REGIONS),
                ("region_type", C.c_int * MAX_REGIONS),
                ("type_temp", C.c_double * MAX_TYPES),
                ("type_seed", C.c_int * MAX_TYPES),
                ("vacuum_rate", C.c_double)]


# name -> argtypes (restype is always c_int = cmd_status unless listed)
_PROTOS = {
    "cmd_last_error": ([], C.c_char_p),
    "cmd_version": ([], C.c_int),
    "cmd_device_init": ([C.c_int], None),
    "cmd_device_count": ([IP], None),
    "cmd_device_synchronize": ([], None),
    "cmd_bandwidth_probe": ([C.c_int64, C.c_int, DP, DP], None),
    "cmd_domain_create": ([DP, DP, C.c_double, C.POINTER(Domain)], None),
    "cmd_system_create": ([C.c_int, C.POINTER(VP)], None),
    "cmd_system_destroy": ([VP], None),
    "cmd_system_set_stream": ([VP, VP], None),
    "cmd_system_set_units": ([VP, C.c_double, C.c_double, C.c_double], None),
    "cmd_system_set_mass": ([VP, C.c_int, DP], None),
    "cmd_system_set_domain": ([VP, C.POINTER(Domain)], None),
    "cmd_system_reserve": ([VP, C.c_int], None),
    "cmd_system_upload": ([VP, C.c_int, C.c_int, DP, DP, DP, IP, IP], None),
    "cmd_system_download": ([VP, DP, DP, DP, IP, IP], None),
    "cmd_system_counts": ([VP, IP, IP, IP], None),
    "cmd_system_footprint": ([VP, LP], None),
    "cmd_nve_initial": ([VP], None),
    "cmd_nve_final": ([VP], None),
    "cmd_comm_exchange": ([VP], None),
    "cmd_comm_exchange_halo": ([VP, C.c_double], None),
    "cmd_comm_update_halo": ([VP], None),
    "cmd_comm_update_force": ([VP], None),
    "cmd_comm_ghost_lists": ([VP, IP, IP], None),
    "cmd_binning_create": ([VP, C.c_double, C.c_double, C.c_double, C.c_int,
                            C.c_int, C.c_int, C.c_int, IP], None),
    "cmd_neigh_create": ([C.c_double, C.c_int, C.c_int, C.POINTER(VP)], None),
    "cmd_neigh_destroy": ([VP], None),
    "cmd_neigh_set_lanes": ([VP, C.c_int], None),
    "cmd_neigh_set_tile": ([VP, C.c_int], None),
    "cmd_neigh_set_tile_width": ([VP, C.c_int], None),
    "cmd_neigh_tile_shape": ([VP, IP, IP], None),
    "cmd_neigh_build": ([VP, VP], None),
    "cmd_neigh_info": ([VP, IP, LP, IP, IP], None),
    "cmd_neigh_layout": ([VP, IP, IP, IP], None),
    "cmd_neigh_download": ([VP, LP, IP], None),
    "cmd_force_lj_create": ([C.c_int, C.POINTER(VP)], None),
    "cmd_force_destroy": ([VP], None),
    "cmd_force_lj_init_coeff": ([VP, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_double], None),
    "cmd_force_set_precision": ([VP, C.c_int], None),
    "cmd_force_set_lanes": ([VP, C.c_int], None),
    "cmd_force_get_lanes": ([VP, IP], None),
    "cmd_force_compute": ([VP, VP, VP], None),
    "cmd_force_energy": ([VP, VP, VP, DP], None),
    "cmd_force_virial": ([VP, VP, VP, DP], None),
    "cmd_nnp_model_create": ([C.c_int, C.c_int, IP, C.c_char_p,
                              C.POINTER(VP)], None),
    "cmd_nnp_model_destroy": ([VP], None),
    "cmd_nnp_model_set_scaling": ([VP, C.c_int, C.c_double, C.c_double], None),
    "cmd_nnp_model_add_sf": ([VP, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_double, C.c_double, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_double], None),
    "cmd_nnp_model_set_layer": ([VP, C.c_int, C.c_int, C.c_int, C.c_int, DP,
                                 DP], None),
    "cmd_nnp_model_set_atom_energy": ([VP, C.c_int, C.c_double], None),
    "cmd_nnp_model_load_n2p2": ([C.c_char_p, C.c_int,
                                 C.POINTER(C.c_char_p), IP, C.POINTER(VP)],
                                None),
    "cmd_nnp_model_synthetic": ([C.c_uint64, C.c_double, C.POINTER(VP)], None),
    "cmd_nnp_model_info": ([VP, C.c_int, IP, IP, IP, DP], None),
    "cmd_nnp_model_get_sf": ([VP, C.c_int, C.c_int, DP], None),
    "cmd_nnp_model_get_layer": ([VP, C.c_int, C.c_int, DP, DP], None),
    "cmd_force_nnp_create": ([VP, C.POINTER(VP)], None),
    "cmd_force_nnp_fit_scaling": ([VP, VP, VP], None),
    "cmd_force_nnp_model": ([VP, C.POINTER(VP)], None),
    "cmd_force_nnp_work": ([VP, C.POINTER(C.c_double)], None),
    "cmd_force_nnp_atomic": ([VP, VP, C.c_int, FP, FP, FP], None),
    "cmd_nnp_lattice_create": ([C.c_int, C.c_int, C.c_int, C.c_double,
                                C.c_double, C.c_double, C.c_uint64, IP, DP,
                                IP, IP], None),
    "cmd_thermo_temperature": ([VP, DP], None),
    "cmd_thermo_kinetic": ([VP, DP], None),
    "cmd_thermo_kinetic_noalloc": ([VP, DP], None),
    "cmd_params_default": ([C.POINTER(Params)], None),
    "cmd_params_parse_deck": ([C.c_char_p, C.c_int, C.POINTER(Params)], None),
    "cmd_lattice_create": ([C.POINTER(Params), C.POINTER(Domain), IP, DP, DP,
                            IP, IP], None),
    "cmd_data_read": ([C.c_char_p, C.c_int, IP, IP, DP, DP, DP, DP, DP, IP,
                       IP, DP], None),
    "cmd_data_write": ([C.c_char_p, C.c_int, C.c_int, DP, DP, DP, DP, IP, IP],
                       None),
    "cmd_md_create": ([C.POINTER(Params), C.POINTER(VP)], None),
    "cmd_md_destroy": ([VP], None),
    "cmd_md_set_precision": ([VP, C.c_int], None),
    "cmd_md_set_lanes": ([VP, C.c_int], None),
    "cmd_md_set_tile": ([VP, C.c_int], None),
    "cmd_md_set_tile_width": ([VP, C.c_int], None),
    "cmd_md_set_ghost_fusion": ([VP, C.c_int], None),
    "cmd_md_set_nnp_model": ([VP, VP], None),
    "cmd_md_set_initial_state": ([VP, C.c_int, DP, DP, IP, IP, DP, DP], None),
    "cmd_md_init": ([VP], None),
    "cmd_md_run": ([VP, C.c_int], None),
    "cmd_md_system": ([VP, C.POINTER(VP)], None),
    "cmd_md_neigh": ([VP, C.POINTER(VP)], None),
    "cmd_md_force": ([VP, C.POINTER(VP)], None),
    "cmd_md_thermo": ([VP, C.c_int, DP, IP], None),
    "cmd_md_timers": ([VP, DP], None),
    "cmd_md_profile": ([VP, C.c_int], None),
    "cmd_md_kernel_stats": ([VP, IP, DP], None),
    "cmd_md_tune_info": ([VP, IP, IP, IP, DP], None),
    "cmd_md_dump": ([VP, C.c_char_p, C.c_int], None),
    "cmd_md_check_correctness": ([VP, C.c_char_p, C.c_int, C.c_char_p, DP],
                                 None),
    "cmd_batch_create": ([C.POINTER(Params), C.POINTER(VP)], None),
    "cmd_batch_destroy": ([VP], None),
    "cmd_batch_set_precision": ([VP, C.c_int], None),
    "cmd_batch_set_lanes": ([VP, C.c_int], None),
    "cmd_batch_set_graphs": ([VP, C.c_int], None),
    "cmd_batch_graph_steps": ([VP, C.POINTER(C.c_longlong)], None),
    "cmd_batch_add": ([VP, C.c_int, DP, DP, IP, IP, DP, DP, IP], None),
    "cmd_batch_remove": ([VP, C.c_int], None),
    "cmd_batch_run": ([VP, C.c_int], None),
    "cmd_batch_counts": ([VP, IP, LP, IP, IP], None),
    "cmd_batch_system_info": ([VP, C.c_int, IP, IP, IP], None),
    "cmd_batch_download": ([VP, C.c_int, DP, DP, DP, IP, IP], None),
    "cmd_batch_thermo": ([VP, C.c_int, C.c_int, DP, IP], None),
    "cmd_batch_timers": ([VP, DP], None),
    "cmd_batch_set_nnp_model": ([VP, VP], None),
    "cmd_server_create": ([C.POINTER(Params), C.POINTER(Params), VP,
                           C.c_int64, C.c_int, C.POINTER(VP)], None),
    "cmd_server_destroy": ([VP], None),
    "cmd_server_configure": ([VP, C.c_int, C.c_int], None),
    "cmd_server_clock": ([VP, DP], None),
    "cmd_server_submit": ([VP, C.c_int, C.c_int, DP, DP, IP, IP, DP, DP,
                           C.c_int, C.c_double, IP], None),
    "cmd_server_poll": ([VP, IP, IP, IP], None),
    "cmd_server_drain": ([VP, C.c_double], None),
    "cmd_server_job": ([VP, C.c_int, IP, DP, IP], None),
    "cmd_server_result": ([VP, C.c_int, DP, DP, IP, C.c_int, DP, IP], None),
    "cmd_server_stats": ([VP, DP], None),
}

_lib = None


def lib():
    """Load the engine library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with python -m cabanamd_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _PROTOS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res if res is not None else C.c_int
        _lib = L
    return _lib


def exported_names():
    return list(_PROTOS)


def check(status):
    if status != CMD_OK:
        msg = lib().cmd_last_error()
        raise CmdError(status, msg.decode() if msg else "")
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def dp(a):
    return a.ctypes.data_as(DP) if a is not None else None


def ip(a):
    return a.ctypes.data_as(IP) if a is not None else None


def lp(a):
    return a.ctypes.data_as(LP) if a is not None else None


def d3(v):
    return (C.c_double * 3)(*[float(t) for t in v])
