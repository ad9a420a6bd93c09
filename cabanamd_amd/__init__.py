"""cabanamd_amd -- MI355X-native CabanaMD timestep engine (Python host side).

The classes mirror the reference's module classes and method names
(CabanaMD src/): System, Comm, Binning, NeighborVerlet, ForceLJ, Integrator,
Temperature/KinE/PotE, InputFile and CbnMD.  Every call goes through the C
ABI of libcabanamd_amd.so (include/cabanamd_amd.h) into hand-written HIP
kernels; nothing here computes physics on the CPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (CmdError, Domain, Params, PREC_FP64, PREC_MIXED, call, d3,
                   dp, ip, lib, lp)

__all__ = ["System", "Comm", "Binning", "NeighborVerlet", "ForceLJ",
           "Integrator", "Temperature", "KinE", "PotE", "InputFile", "CbnMD",
           "CmdError", "Domain", "Params", "PREC_FP64", "PREC_MIXED",
           "create_domain", "device_count", "lattice_params"]


def device_count():
    c = C.c_int(0)
    call("cmd_device_count", C.byref(c))
    return c.value


def create_domain(low, high, ghost_cutoff):
    """System::create_domain (system.h:149-205) as a plain struct."""
    d = Domain()
    call("cmd_domain_create", d3(low), d3(high), C.c_double(ghost_cutoff),
         C.byref(d))
    return d


class System:
    """System<Device, layout> (system.h:65-279): device-resident atoms."""

    def __init__(self, ntypes=1):
        h = C.c_void_p()
        call("cmd_system_create", ntypes, C.byref(h))
        self._h = h
        self.ntypes = ntypes
        self.mass = np.ones(ntypes)
        self.dt, self.mvv2e, self.boltz = 0.005, 1.0, 1.0
        self.domain = None

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cmd_system_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_units(self, dt, mvv2e=1.0, boltz=1.0):
        self.dt, self.mvv2e, self.boltz = dt, mvv2e, boltz
        call("cmd_system_set_units", self._h, dt, mvv2e, boltz)

    def set_mass(self, mass):
        m = np.ascontiguousarray(mass, dtype=np.float64)
        self.mass = m
        call("cmd_system_set_mass", self._h, len(m), dp(m))

    def create_domain(self, low, high, ghost_cutoff=None):
        if ghost_cutoff is None:
            ghost_cutoff = max(high[0] - low[0], high[2] - low[1],
                               high[2] - low[2])
        self.domain = create_domain(low, high, ghost_cutoff)
        call("cmd_system_set_domain", self._h, C.byref(self.domain))
        return self.domain

    def set_domain(self, d):
        self.domain = d
        call("cmd_system_set_domain", self._h, C.byref(d))

    def resize(self, n):
        call("cmd_system_reserve", self._h, int(n))

    def upload(self, x, v=None, f=None, type=None, id=None, n_ghost=0):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        t = (np.zeros(n, dtype=np.int32) if type is None
             else np.ascontiguousarray(type, dtype=np.int32))
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        f = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        i = None if id is None else np.ascontiguousarray(id, dtype=np.int32)
        call("cmd_system_upload", self._h, n - n_ghost, n_ghost, dp(x), dp(v),
             dp(f), ip(t), ip(i))

    def counts(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("cmd_system_counts", self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    @property
    def N(self):
        return self.counts()[0]

    @property
    def N_local(self):
        return self.counts()[1]

    @property
    def N_ghost(self):
        return self.counts()[2]

    def download(self):
        """(x, v, f, type, id) of local + ghost atoms as host arrays."""
        _, nl, ng = self.counts()
        n = nl + ng
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 3))
        t = np.zeros(n, dtype=np.int32)
        i = np.zeros(n, dtype=np.int32)
        if n:
            call("cmd_system_download", self._h, dp(x), dp(v), dp(f), ip(t),
                 ip(i))
        return x, v, f, t, i

    # slice_* of the reference return views; here they return host copies
    def slice_x(self):
        return self.download()[0]

    def slice_v(self):
        return self.download()[1]

    def slice_f(self):
        return self.download()[2]

    def slice_type(self):
        return self.download()[3]

    def slice_id(self):
        return self.download()[4]

    def footprint(self):
        b = C.c_int64()
        call("cmd_system_footprint", self._h, C.byref(b))
        return b.value


class Integrator:
    """Integrator<System> (integrator_nve.h:56-110): velocity Verlet NVE."""

    def __init__(self, system):
        self.system = system

    def initial_integrate(self, system=None):
        call("cmd_nve_initial", (system or self.system).handle)

    def final_integrate(self, system=None):
        call("cmd_nve_final", (system or self.system).handle)

    def name(self):
        return "Integrator:NVE"


class Comm:
    """Comm<System> on one rank (comm_mpi_impl.h:191-441)."""

    def __init__(self, system, comm_depth):
        self.system = system
        self.comm_depth = comm_depth

    def exchange(self):
        call("cmd_comm_exchange", self.system.handle)

    def exchange_halo(self):
        call("cmd_comm_exchange_halo", self.system.handle, self.comm_depth)

    def update_halo(self):
        call("cmd_comm_update_halo", self.system.handle)

    def update_force(self):
        call("cmd_comm_update_force", self.system.handle)

    def ghost_lists(self):
        counts = np.zeros(6, dtype=np.int32)
        call("cmd_comm_ghost_lists", self.system.handle, None, ip(counts))
        send = np.zeros(max(int(counts.sum()), 1), dtype=np.int32)
        call("cmd_comm_ghost_lists", self.system.handle, ip(send), ip(counts))
        return send[:counts.sum()], counts

    def name(self):
        return "Comm:HIP"


class Binning:
    """Binning<System> (binning_cabana_impl.h:57-113)."""

    def __init__(self, system):
        self.system = system

    def create_binning(self, dx, dy, dz, halo_depth=1, do_local=True,
                       do_ghost=False, sort=True):
        n = self.system.N_local
        perm = np.zeros(max(n, 1), dtype=np.int32)
        call("cmd_binning_create", self.system.handle, dx, dy, dz, halo_depth,
             int(do_local), int(do_ghost), int(sort), ip(perm))
        return perm[:n]

    def name(self):
        return "Binning:HIPLinkedCell"


class NeighborVerlet:
    """NeighborVerlet<System, Full|Half, CSR> (neighbor_verlet.h:19-70)."""

    def __init__(self, neigh_cut, half_neigh, max_neigh_guess=50, lanes=1,
                 tile=0, tile_width=0):
        h = C.c_void_p()
        call("cmd_neigh_create", neigh_cut, int(half_neigh),
             int(max_neigh_guess), C.byref(h))
        self._h = h
        call("cmd_neigh_set_lanes", h, int(lanes))
        call("cmd_neigh_set_tile", h, int(tile))
        call("cmd_neigh_set_tile_width", h, int(tile_width))
        self.neigh_cut = neigh_cut
        self.half_neigh = bool(half_neigh)

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self, "_own", True):
            lib().cmd_neigh_destroy(self._h)
            self._h = None

    @classmethod
    def _borrow(cls, handle, neigh_cut, half):
        o = cls.__new__(cls)
        o._h = handle
        o._own = False
        o.neigh_cut = neigh_cut
        o.half_neigh = bool(half)
        return o

    @property
    def handle(self):
        return self._h

    def create(self, system):
        call("cmd_neigh_build", self._h, system.handle)

    def info(self):
        r, nnz, mr, g = C.c_int(), C.c_int64(), C.c_int(), C.c_int()
        call("cmd_neigh_info", self._h, C.byref(r), C.byref(nnz), C.byref(mr),
             C.byref(g))
        return dict(n_rows=r.value, nnz=nnz.value, max_row=mr.value,
                    max_neigh_guess=g.value)

    def layout(self):
        """Layout of the last build: tiled (bool), tile height, lanes."""
        t, e, g, w = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        call("cmd_neigh_layout", self._h, C.byref(t), C.byref(e), C.byref(g))
        call("cmd_neigh_tile_shape", self._h, C.byref(w), None)
        return dict(tiled=bool(t.value), tile_height=e.value, lanes=g.value,
                    tile_width=w.value)

    @property
    def max_neigh_guess(self):
        return self.info()["max_neigh_guess"]

    def get(self):
        """CSR list (offsets[n_rows+1] int64, idx[nnz] int32) on the host."""
        inf = self.info()
        off = np.zeros(inf["n_rows"] + 1, dtype=np.int64)
        idx = np.zeros(max(inf["nnz"], 1), dtype=np.int32)
        call("cmd_neigh_download", self._h, lp(off), ip(idx))
        return off, idx[:inf["nnz"]]

    def name(self):
        return ("Neighbor:HIPVerletHalf" if self.half_neigh
                else "Neighbor:HIPVerletFull")


class NNPModel:
    """The model ForceNNP::init_coeff loads (force_nnp_cabana_neigh_impl.h:
    41-66): symmetry functions, scaling ranges and per-element networks."""

    def __init__(self, handle, own=True, owner=None):
        self._h = handle
        self._own = own
        self._owner = owner  # keeps the object that owns a borrowed model

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self, "_own", True):
            lib().cmd_nnp_model_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @classmethod
    def _borrowed(cls, handle, owner):
        return cls(handle, own=False, owner=owner)

    @classmethod
    def create(cls, nelem, hidden, activations):
        h = C.c_void_p()
        hid = (C.c_int * max(len(hidden), 1))(*hidden)
        call("cmd_nnp_model_create", nelem, len(hidden), hid,
             activations.encode(), C.byref(h))
        return cls(h)

    @classmethod
    def load_n2p2(cls, path, elements, atomic_numbers):
        h = C.c_void_p()
        names = (C.c_char_p * len(elements))(*[e.encode() for e in elements])
        z = (C.c_int * len(elements))(*atomic_numbers)
        call("cmd_nnp_model_load_n2p2", str(path).encode(), len(elements),
             names, z, C.byref(h))
        return cls(h)

    @classmethod
    def synthetic(cls, seed=87287, rc=6.0):
        h = C.c_void_p()
        call("cmd_nnp_model_synthetic", C.c_uint64(seed), rc, C.byref(h))
        return cls(h)

    def set_scaling(self, scale=True, smin=0.0, smax=1.0):
        call("cmd_nnp_model_set_scaling", self._h, int(scale), smin, smax)

    def add_sf(self, elem, typ, e1, e2, eta, rs, lam, zeta, rc, gmin=0.0,
               gmax=1.0):
        call("cmd_nnp_model_add_sf", self._h, elem, typ, e1, e2, eta, rs,
             lam, zeta, rc, gmin, gmax)

    def set_layer(self, elem, layer, W, b):
        W = np.ascontiguousarray(W, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        call("cmd_nnp_model_set_layer", self._h, elem, layer, W.shape[0],
             W.shape[1], dp(W), dp(b))

    def set_atom_energy(self, elem, e):
        call("cmd_nnp_model_set_atom_energy", self._h, elem, e)

    def describe(self, nelem):
        """The model as a plain dict (the layout of tests/nnp_models.py)."""
        out = dict(nelem=nelem, sf=[], gmin=[], gmax=[], W=[], b=[],
                   atom_energy=[])
        for e in range(nelem):
            ns, nh = C.c_int(), C.c_int()
            hid = (C.c_int * 8)()
            ae = C.c_double()
            call("cmd_nnp_model_info", self._h, e, C.byref(ns), C.byref(nh),
                 hid, C.byref(ae))
            hidden = list(hid[:nh.value])
            out["hidden"] = hidden
            sf, gmin, gmax = [], [], []
            prm = (C.c_double * 10)()
            for k in range(ns.value):
                call("cmd_nnp_model_get_sf", self._h, e, k, prm)
                v = list(prm)
                sf.append((int(v[0]), int(v[1]), int(v[2]), v[3], v[4], v[5],
                           v[6], v[7]))
                gmin.append(v[8])
                gmax.append(v[9])
            sizes = [ns.value] + hidden + [1]
            Ws, bs = [], []
            for l in range(len(sizes) - 1):
                W = np.zeros((sizes[l + 1], sizes[l]))
                b = np.zeros(sizes[l + 1])
                call("cmd_nnp_model_get_layer", self._h, e, l, dp(W), dp(b))
                Ws.append(W)
                bs.append(b)
            out["sf"].append(sf)
            out["gmin"].append(np.array(gmin))
            out["gmax"].append(np.array(gmax))
            out["W"].append(Ws)
            out["b"].append(bs)
            out["atom_energy"].append(ae.value)
        return out


class ForceNNP:
    """ForceNNP (force_types/force_nnp_cabana_neigh_impl.h:41-130)."""

    def __init__(self, model, precision=PREC_MIXED):
        h = C.c_void_p()
        call("cmd_force_nnp_create", model.handle, C.byref(h))
        self._h = h
        call("cmd_force_set_precision", h, precision)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cmd_force_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_precision(self, prec):
        call("cmd_force_set_precision", self._h, prec)

    def fit_scaling(self, system, neighbor):
        call("cmd_force_nnp_fit_scaling", self._h, system.handle,
             neighbor.handle)

    def model(self):
        m = C.c_void_p()
        call("cmd_force_nnp_model", self._h, C.byref(m))
        # a view into this force's model: it keeps the force alive
        return NNPModel._borrowed(m, self)

    def compute(self, system, neighbor):
        call("cmd_force_compute", self._h, system.handle, neighbor.handle)

    def compute_energy(self, system, neighbor):
        e = C.c_double()
        call("cmd_force_energy", self._h, system.handle, neighbor.handle,
             C.byref(e))
        return e.value

    def atomic(self, system, ld):
        n = system.N_local
        G = np.zeros((n, ld), dtype=np.float32)
        D = np.zeros((n, ld), dtype=np.float32)
        E = np.zeros(n, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        call("cmd_force_nnp_atomic", self._h, system.handle, ld,
             G.ctypes.data_as(fp), D.ctypes.data_as(fp), E.ctypes.data_as(fp))
        return G, D, E

    def work(self):
        """Work model of the last compute: neighbours and triplets summed
        over the atoms, FP32 flops of the symmetry functions, the force chain
        and the layers, and the atom count."""
        w = (C.c_double * 6)()
        call("cmd_force_nnp_work", self._h, w)
        return dict(neighbours=int(w[0]), triplets=int(w[1]),
                    flop_symmetry_functions=w[2], flop_force_chain=w[3],
                    flop_layers=w[4], atoms=int(w[5]))

    def name(self):
        return "Force:NNPHIP"


def nnp_lattice(cells, spacing=2.5, jitter=0.1, frac1=1.0 / 3.0, seed=87287):
    """Jittered simple-cubic sites with a seeded 2:1 element split."""
    n = C.c_int()
    call("cmd_nnp_lattice_create", cells[0], cells[1], cells[2], spacing,
         jitter, frac1, C.c_uint64(seed), C.byref(n), None, None, None)
    x = np.zeros((n.value, 3))
    t = np.zeros(n.value, dtype=np.int32)
    i = np.zeros(n.value, dtype=np.int32)
    call("cmd_nnp_lattice_create", cells[0], cells[1], cells[2], spacing,
         jitter, frac1, C.c_uint64(seed), C.byref(n), dp(x), ip(t), ip(i))
    return x, t, i


class ForceLJ:
    """ForceLJ (force_types/force_lj_cabana_neigh_impl.h:49-377)."""

    def __init__(self, system, precision=PREC_FP64):
        h = C.c_void_p()
        call("cmd_force_lj_create", system.ntypes, C.byref(h))
        self._h = h
        self._own = True
        call("cmd_force_set_precision", h, precision)

    @classmethod
    def _borrow(cls, handle):
        o = cls.__new__(cls)
        o._h = handle
        o._own = False
        return o

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self, "_own", True):
            lib().cmd_force_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def init_coeff(self, args):
        """args: pair_coeff lines as word lists, e.g.
        [["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]]."""
        for a in args:
            call("cmd_force_lj_init_coeff", self._h, int(a[1]), int(a[2]),
                 float(a[3]), float(a[4]), float(a[5]))

    def set_precision(self, prec):
        call("cmd_force_set_precision", self._h, prec)

    def set_lanes(self, lanes):
        call("cmd_force_set_lanes", self._h, lanes)

    @property
    def lanes(self):
        g = C.c_int()
        call("cmd_force_get_lanes", self._h, C.byref(g))
        return g.value

    def compute(self, system, neighbor):
        call("cmd_force_compute", self._h, system.handle, neighbor.handle)

    def compute_energy(self, system, neighbor):
        e = C.c_double()
        call("cmd_force_energy", self._h, system.handle, neighbor.handle,
             C.byref(e))
        return e.value

    def compute_virial(self, system, neighbor):
        w = C.c_double()
        call("cmd_force_virial", self._h, system.handle, neighbor.handle,
             C.byref(w))
        return w.value

    def name(self):
        return "Force:LJHIP"


class Temperature:
    """Temperature (property_temperature_impl.h:55-77)."""

    def __init__(self, comm=None):
        self.comm = comm

    def compute(self, system):
        t = C.c_double()
        call("cmd_thermo_temperature", system.handle, C.byref(t))
        return t.value


class KinE:
    """KinE (property_kine_impl.h:55-76)."""

    def __init__(self, comm=None):
        self.comm = comm

    def compute(self, system):
        k = C.c_double()
        call("cmd_thermo_kinetic", system.handle, C.byref(k))
        return k.value


class PotE:
    """PotE (property_pote_impl.h:55-63)."""

    def __init__(self, comm=None):
        self.comm = comm

    def compute(self, system, force, neighbor):
        return force.compute_energy(system, neighbor)


class InputFile:
    """InputFile<System>::read_file (inputFile_impl.h:86-534)."""

    def __init__(self, force_iteration=None):
        self.force_iteration = force_iteration
        self.params = Params()
        call("cmd_params_default", C.byref(self.params))

    def read_file(self, path):
        fi = -1 if self.force_iteration is None else int(self.force_iteration)
        call("cmd_params_parse_deck", str(path).encode(), fi,
             C.byref(self.params))
        return self.params

    def create_lattice(self, domain):
        """Host-side lattice + velocities (inputFile_impl.h:536-865)."""
        p = self.params
        n = C.c_int()
        call("cmd_lattice_create", C.byref(p), C.byref(domain), C.byref(n),
             None, None, None, None)
        x = np.zeros((n.value, 3))
        v = np.zeros((n.value, 3))
        t = np.zeros(n.value, dtype=np.int32)
        i = np.zeros(n.value, dtype=np.int32)
        call("cmd_lattice_create", C.byref(p), C.byref(domain), C.byref(n),
             dp(x), dp(v), ip(t), ip(i))
        return x, v, t, i


def nnp_params(ntypes=2, cutoff=6.0, skin=0.3, dt=0.0005, nsteps=200,
               thermo=20, every=20, mass=(1.0, 1.0), max_neigh_guess=100):
    """Run parameters for an NNP system whose atoms come from
    CbnMD.set_initial_state and whose model from CbnMD.set_nnp_model
    (pair_style nnp, pair_coeff * * cutoff; reduced units)."""
    p = Params()
    call("cmd_params_default", C.byref(p))
    p.units = 1
    p.ntypes = ntypes
    for k in range(ntypes):
        p.mass[k] = mass[k]
    p.force_type = 2
    p.force_cutoff = cutoff
    p.skin = skin
    p.every = every
    p.max_neigh_guess = max_neigh_guess
    p.comm_ghost_cutoff = cutoff + skin
    p.half = 0
    p.nsteps = nsteps
    p.thermo_rate = thermo
    p.dt = dt
    p.mvv2e = 1.0
    p.boltz = 1.0
    return p


def lattice_params(ncells=(20, 20, 20), rho=0.8442, temp=1.44, seed=87287,
                   mass=1.0, cutoff=2.5, skin=0.3, every=20, half=False,
                   nsteps=100, thermo=10, dt=0.005, gaussian=False,
                   ghost_cutoff=None):
    """Run parameters of the fcc LJ workload (an in.lj-style deck)."""
    p = Params()
    call("cmd_params_default", C.byref(p))
    p.units = 1
    p.lattice_style = 1
    p.lattice_constant = (4.0 / rho) ** (1.0 / 3.0)
    for k in range(3):
        p.region_lo[k] = 0.0
        p.region_hi[k] = float(ncells[k])
    p.ntypes = 1
    p.mass[0] = mass
    p.temp_target = temp
    p.temp_seed = seed
    p.vel_gaussian = 1 if gaussian else 0
    p.force_cutoff = cutoff
    p.skin = skin
    p.every = every
    if ghost_cutoff is not None:
        p.comm_ghost_cutoff = ghost_cutoff
    p.half = 1 if half else 0
    p.nsteps = nsteps
    p.thermo_rate = thermo
    p.dt = dt
    p.mvv2e = 1.0
    p.boltz = 1.0
    p.ncoeff = 1
    p.coeff_i[0] = 1
    p.coeff_j[0] = 1
    p.coeff_eps[0] = 1.0
    p.coeff_sigma[0] = 1.0
    p.coeff_cut[0] = cutoff
    p.force_type = 0
    return p


class CbnMD:
    """CbnMD<System, Neighbor> (cabanamd_impl.h:70-432): init() then run()."""

    def __init__(self, params, precision=PREC_FP64, lanes=0, tile=0,
                 fuse_ghosts=True, tile_width=0):
        self.params = params
        h = C.c_void_p()
        call("cmd_md_create", C.byref(params), C.byref(h))
        self._h = h
        call("cmd_md_set_precision", h, precision)
        call("cmd_md_set_lanes", h, lanes)
        call("cmd_md_set_tile", h, tile)
        call("cmd_md_set_tile_width", h, tile_width)
        call("cmd_md_set_ghost_fusion", h, int(bool(fuse_ghosts)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cmd_md_destroy(self._h)
            self._h = None

    def set_nnp_model(self, model):
        """Use this NNP model instead of the deck's pair_style nnp dir."""
        call("cmd_md_set_nnp_model", self._h, model.handle)

    def set_initial_state(self, x, type, lo, hi, v=None, id=None):
        """Start from these atoms in the box [lo, hi) instead of the deck's
        lattice and velocities."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        t = np.ascontiguousarray(type, dtype=np.int32)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        ii = None if id is None else np.ascontiguousarray(id, dtype=np.int32)
        call("cmd_md_set_initial_state", self._h, n, dp(x),
             None if vv is None else dp(vv), ip(t),
             None if ii is None else ip(ii), d3(lo), d3(hi))

    def init(self):
        call("cmd_md_init", self._h)
        s = C.c_void_p()
        call("cmd_md_system", self._h, C.byref(s))
        self._sys = s
        n = C.c_void_p()
        call("cmd_md_neigh", self._h, C.byref(n))
        self.neighbor = NeighborVerlet._borrow(
            n, self.params.force_cutoff + self.params.skin, self.params.half)
        fh = C.c_void_p()
        call("cmd_md_force", self._h, C.byref(fh))
        self.force = ForceLJ._borrow(fh)

    def run(self, nsteps=None):
        call("cmd_md_run", self._h,
             int(self.params.nsteps if nsteps is None else nsteps))

    def dump(self, path, step):
        """dump_binary (cabanamd_impl.h:434-481): PATH/output.<step>.000."""
        call("cmd_md_dump", self._h, str(path).encode(), int(step))

    def check_correctness(self, ref_path, step, file=None):
        """check_correctness (cabanamd_impl.h:485-642) against the snapshot
        of `step`: |dx|, max dx, |dv|, max dv, |df|, max df."""
        out = np.zeros(6)
        call("cmd_md_check_correctness", self._h, str(ref_path).encode(),
             int(step), None if file is None else str(file).encode(), dp(out))
        return out

    def compute_forces(self):
        """Force::compute on the current positions with the current list
        (and the ghost-force fold for a half list), as at a rebuild-free
        step: leaves the per-atom forces readable by arrays()."""
        call("cmd_force_compute", self.force.handle, self._sys,
             self.neighbor.handle)
        if self.params.half or self.params.force_type == 2:
            call("cmd_comm_update_force", self._sys)

    def arrays(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("cmd_system_counts", self._sys, C.byref(a), C.byref(b),
             C.byref(c))
        n = b.value + c.value
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 3))
        t = np.zeros(n, dtype=np.int32)
        i = np.zeros(n, dtype=np.int32)
        call("cmd_system_download", self._sys, dp(x), dp(v), dp(f), ip(t),
             ip(i))
        return x, v, f, t, i

    @property
    def n_local(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("cmd_system_counts", self._sys, C.byref(a), C.byref(b),
             C.byref(c))
        return b.value

    def thermo(self):
        """Rows (step, T, PE/N, (PE+KE)/N, virial)."""
        n = C.c_int()
        call("cmd_md_thermo", self._h, 0, None, C.byref(n))
        rows = np.zeros((max(n.value, 1), 5))
        call("cmd_md_thermo", self._h, n.value, dp(rows), C.byref(n))
        return rows[:n.value]

    def timers(self):
        t = np.zeros(6)
        call("cmd_md_timers", self._h, dp(t))
        return dict(zip(["force", "neigh", "comm", "int", "other", "total"],
                        t))

    def profile(self, enable=True):
        call("cmd_md_profile", self._h, int(enable))

    def kernel_stats(self):
        n, ms = C.c_int(), C.c_double()
        call("cmd_md_kernel_stats", self._h, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def tune_info(self):
        """(lanes per atom, tile height or 0, seconds init spent choosing,
        tile width or 0)."""
        g, w, e, s = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        call("cmd_md_tune_info", self._h, C.byref(g), C.byref(w), C.byref(e),
             C.byref(s))
        return g.value, e.value, s.value, w.value


class Batch:
    """Many independent LJ systems stepped together (no reference
    counterpart; each system follows CbnMD::run, cabanamd_impl.h:285-399).
    params: shared run parameters (pair_coeff, cutoff, skin, every, dt,
    thermo_rate); systems join with add() and leave with remove() at the
    first run() that starts on a rebuild step."""

    def __init__(self, params, precision=PREC_FP64, lanes=0, nnp_model=None,
                 graphs=True):
        self.params = params
        h = C.c_void_p()
        call("cmd_batch_create", C.byref(params), C.byref(h))
        self._h = h
        call("cmd_batch_set_precision", h, precision)
        call("cmd_batch_set_lanes", h, lanes)
        call("cmd_batch_set_graphs", h, int(bool(graphs)))
        self._model = nnp_model
        if nnp_model is not None:
            call("cmd_batch_set_nnp_model", h, nnp_model.handle)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cmd_batch_destroy(self._h)
            self._h = None

    def add(self, x, type, lo, hi, v=None, id=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        t = np.ascontiguousarray(type, dtype=np.int32)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        ii = None if id is None else np.ascontiguousarray(id, dtype=np.int32)
        h = C.c_int()
        call("cmd_batch_add", self._h, n, dp(x),
             None if vv is None else dp(vv), ip(t),
             None if ii is None else ip(ii), d3(lo), d3(hi), C.byref(h))
        return h.value

    def remove(self, handle):
        call("cmd_batch_remove", self._h, int(handle))

    def set_precision(self, prec):
        call("cmd_batch_set_precision", self._h, int(prec))

    def run(self, nsteps):
        call("cmd_batch_run", self._h, int(nsteps))

    def graph_steps(self):
        """Steps run by replays of the captured plain-step graph."""
        n = C.c_longlong()
        call("cmd_batch_graph_steps", self._h, C.byref(n))
        return n.value

    def counts(self):
        a, c, d = C.c_int(), C.c_int(), C.c_int()
        b = C.c_int64()
        call("cmd_batch_counts", self._h, C.byref(a), C.byref(b), C.byref(c),
             C.byref(d))
        return dict(systems=a.value, atoms=b.value, step=c.value,
                    lanes=d.value)

    def info(self, handle):
        n, s, st = C.c_int(), C.c_int(), C.c_int()
        call("cmd_batch_system_info", self._h, int(handle), C.byref(n),
             C.byref(s), C.byref(st))
        return dict(n=n.value, step=s.value,
                    state=("unknown", "waiting", "running", "leaving")[st.value])

    def download(self, handle, by_id=True):
        """x, v, f, type, id of one system (sorted by id when by_id)."""
        n = self.info(handle)["n"]
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 3))
        t = np.zeros(n, dtype=np.int32)
        i = np.zeros(n, dtype=np.int32)
        call("cmd_batch_download", self._h, int(handle), dp(x), dp(v), dp(f),
             ip(t), ip(i))
        if by_id:
            o = np.argsort(i, kind="stable")
            x, v, f, t, i = x[o], v[o], f[o], t[o], i[o]
        return x, v, f, t, i

    def thermo(self, handle):
        """Rows (local step, T, PE/N, (PE+KE)/N)."""
        r = C.c_int()
        call("cmd_batch_thermo", self._h, int(handle), 0, None, C.byref(r))
        out = np.zeros((r.value, 4))
        call("cmd_batch_thermo", self._h, int(handle), r.value, dp(out),
             C.byref(r))
        return out

    def timers(self):
        t = np.zeros(4)
        call("cmd_batch_timers", self._h, dp(t))
        return dict(rebuild=t[0], steps=t[1], thermo=t[2], join=t[3])


def fcc_system(ncells, rho, temp, seed, mass=1.0, gaussian=True):
    """An fcc LJ box of 4*nx*ny*nz atoms at reduced density rho with
    velocities at temperature temp (the lattice and velocity code of
    InputFile::create_lattice, inputFile_impl.h:536-865).  Returns x, v,
    type, id, lo, hi."""
    p = lattice_params(ncells=ncells, rho=rho, temp=temp, seed=seed,
                       mass=mass, gaussian=gaussian)
    a = p.lattice_constant
    hi = [a * c for c in ncells]
    d = create_domain([0.0, 0.0, 0.0], hi, p.force_cutoff + p.skin)
    inp = InputFile()
    inp.params = p
    x, v, t, i = inp.create_lattice(d)
    return x, v, t, i, [0.0, 0.0, 0.0], hi


def nnp_system(cells, spacing=2.5, jitter=0.1, temp=0.05, seed=87287,
               mass=1.0):
    """A two-element NNP box: jittered simple-cubic lattice (one third
    element 1) with Gaussian velocities at temperature temp and zero net
    momentum.  Returns x, v, type, id, lo, hi."""
    x, t, i = nnp_lattice(cells, spacing=spacing, jitter=jitter, seed=seed)
    rng = np.random.default_rng(seed + 17)
    v = rng.normal(0.0, np.sqrt(temp / mass), x.shape)
    v -= v.mean(axis=0)
    hi = [spacing * c for c in cells]
    x = np.mod(x, np.array(hi))
    return x, v, t, i, [0.0, 0.0, 0.0], hi


class Server:
    """MD request server with continuous batching (no reference
    counterpart): LJ jobs and, given nnp params and a model, NNP jobs, each
    kind in one device-resident batch; jobs are admitted in arrival order by
    device footprint and slot, join at rebuild boundaries and leave with
    their results when they finish."""

    LJ, NNP = 0, 2
    QUEUED, RUNNING, DONE = 1, 2, 3

    def __init__(self, lj_params, nnp_params=None, nnp_model=None,
                 hbm_budget=64 << 30, max_jobs=64, precision=PREC_MIXED,
                 lanes=0):
        h = C.c_void_p()
        call("cmd_server_create", C.byref(lj_params),
             None if nnp_params is None else C.byref(nnp_params),
             None if nnp_model is None else nnp_model.handle,
             C.c_int64(int(hbm_budget)), int(max_jobs), C.byref(h))
        self._h = h
        self._model = nnp_model
        self._n = {}
        call("cmd_server_configure", h, precision, lanes)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cmd_server_destroy(self._h)
            self._h = None

    def clock(self):
        t = C.c_double()
        call("cmd_server_clock", self._h, C.byref(t))
        return t.value

    def submit(self, kind, x, type, lo, hi, nsteps, v=None, id=None,
               arrival=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        t = np.ascontiguousarray(type, dtype=np.int32)
        vv = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        ii = None if id is None else np.ascontiguousarray(id, dtype=np.int32)
        j = C.c_int()
        call("cmd_server_submit", self._h, int(kind), n, dp(x),
             None if vv is None else dp(vv), ip(t),
             None if ii is None else ip(ii), d3(lo), d3(hi), int(nsteps),
             float(arrival), C.byref(j))
        self._n[j.value] = n
        return j.value

    def poll(self):
        q, r, d = C.c_int(), C.c_int(), C.c_int()
        call("cmd_server_poll", self._h, C.byref(q), C.byref(r), C.byref(d))
        return q.value, r.value, d.value

    def drain(self, timeout=600.0):
        call("cmd_server_drain", self._h, float(timeout))

    def job(self, j):
        st, n = C.c_int(), C.c_int()
        t = np.zeros(3)
        call("cmd_server_job", self._h, int(j), C.byref(st), dp(t), C.byref(n))
        return dict(state=st.value, arrival=t[0], admit=t[1], done=t[2],
                    n=n.value)

    def result(self, j):
        """x, v, id (ordered by id) and thermo rows of a finished job."""
        n = self._n[j]
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        i = np.zeros(n, dtype=np.int32)
        r = C.c_int()
        call("cmd_server_result", self._h, int(j), dp(x), dp(v), ip(i), 0,
             None, C.byref(r))
        th = np.zeros((r.value, 4))
        call("cmd_server_result", self._h, int(j), None, None, None, r.value,
             dp(th), C.byref(r))
        return x, v, i, th

    def stats(self):
        o = np.zeros(6)
        call("cmd_server_stats", self._h, dp(o))
        return dict(atom_steps=o[0], busy_s=o[1], rounds=int(o[2]),
                    bytes=int(o[3]), running=int(o[4]), done=int(o[5]))


def read_data(path, charge=False, ntypes_mass=8):
    """A LAMMPS data file (read_data.h:100-385): dict with x, v, q, type
    (0-based), id, lo, hi, ntypes and mass (NaN where the file has none)."""
    n, nt = C.c_int(), C.c_int()
    lo, hi = np.zeros(3), np.zeros(3)
    call("cmd_data_read", str(path).encode(), int(bool(charge)), C.byref(n),
         C.byref(nt), dp(lo), dp(hi), None, None, None, None, None, None)
    k = n.value
    x, v, q = np.zeros((k, 3)), np.zeros((k, 3)), np.zeros(k)
    t, i = np.zeros(k, np.int32), np.zeros(k, np.int32)
    mass = np.full(ntypes_mass, np.nan)
    call("cmd_data_read", str(path).encode(), int(bool(charge)), C.byref(n),
         C.byref(nt), dp(lo), dp(hi), dp(x), dp(v), dp(q), ip(t), ip(i),
         dp(mass))
    return dict(x=x, v=v, q=q, type=t, id=i, lo=lo, hi=hi, ntypes=nt.value,
                mass=mass[:nt.value])


def write_data(path, x, v, type, id, lo, hi, ntypes):
    """write_data (read_data.h:387-422) at full double precision."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(type, dtype=np.int32)
    i = np.ascontiguousarray(id, dtype=np.int32)
    call("cmd_data_write", str(path).encode(), len(x), int(ntypes),
         dp(np.asarray(lo, dtype=np.float64)),
         dp(np.asarray(hi, dtype=np.float64)), dp(x), dp(v), ip(t), ip(i))

