"""ctypes view of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package
(cabanamd_amd) never imports this module.

Each wrapper mirrors one function of oracle.h, which cites the CabanaMD
reference lines it restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)
LP = C.POINTER(C.c_int64)


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
                ("ntypes", C.c_int), ("mass", C.c_double * 8),
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
                ("nregions", C.c_int),
                ("region_name", (C.c_char * 32) * 16),
                ("region_block", (C.c_double * 6) * 16),
                ("region_type", C.c_int * 16),
                ("type_temp", C.c_double * 8),
                ("type_seed", C.c_int * 8)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(path)
        L.orc_rng_uniform.restype = C.c_double
        L.orc_rng_gaussian.restype = C.c_double
        L.orc_lj_energy.restype = C.c_double
        L.orc_lj_virial.restype = C.c_double
        L.orc_temperature.restype = C.c_double
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_neigh_build.restype = C.c_int64
        L.orc_neigh_brute.restype = C.c_int64
        L.orc_neigh_build_grid.restype = C.c_int64
        L.orc_md_create.restype = C.c_void_p
        L.orc_md_destroy.argtypes = [C.c_void_p]
        L.orc_md_init.argtypes = [C.c_void_p]
        L.orc_md_run.argtypes = [C.c_void_p, C.c_int]
        L.orc_md_n_local.argtypes = [C.c_void_p]
        L.orc_md_n_ghost.argtypes = [C.c_void_p]
        L.orc_md_nnz.argtypes = [C.c_void_p]
        L.orc_md_nnz.restype = C.c_int64
        L.orc_md_get.argtypes = [C.c_void_p, DP, DP, DP, IP, IP]
        L.orc_md_thermo.argtypes = [C.c_void_p, C.c_int, DP]
        L.orc_md_timers.argtypes = [C.c_void_p, DP]
        _LIB = L
    return _LIB


def dp(a):
    return a.ctypes.data_as(DP) if a is not None else None


def ip(a):
    return a.ctypes.data_as(IP) if a is not None else None


def lp(a):
    return a.ctypes.data_as(LP) if a is not None else None


def _d3(v):
    return (C.c_double * 3)(*[float(t) for t in v])


def lj_params(ncells=(20, 20, 20), rho=0.8442, temp=1.44, seed=87287,
              mass=1.0, cutoff=2.5, skin=0.3, every=20, half=False,
              nsteps=100, thermo=10, dt=0.005, gaussian=False,
              ghost_cutoff=None):
    """Parameters of the fcc LJ workload (in.lj-style deck, lj units)."""
    p = Params()
    p.units = 1
    p.lattice_style = 1
    p.lattice_constant = (4.0 / rho) ** (1.0 / 3.0)
    p.lattice_offset = _d3((0, 0, 0))
    p.region_lo = _d3((0, 0, 0))
    p.region_hi = _d3(ncells)
    p.ntypes = 1
    p.mass[0] = mass
    p.temp_target = temp
    p.temp_seed = seed
    p.vel_gaussian = 1 if gaussian else 0
    p.force_cutoff = cutoff
    p.skin = skin
    p.every = every
    p.max_neigh_guess = 50
    # default comm_ghost_cutoff: 20 lattice constants at the default
    # lattice_constant 0.8442 (inputFile_impl.h:82-83)
    p.comm_ghost_cutoff = (ghost_cutoff if ghost_cutoff is not None
                           else (4.0 / 0.8442) ** (1.0 / 3.0) * 20.0)
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
    return p


def create_domain(low, high, ghost_cutoff):
    d = Domain()
    lib().orc_create_domain(_d3(low), _d3(high), C.c_double(ghost_cutoff),
                            C.byref(d))
    return d


def crand(n):
    """The first n values of the C library's rand() after srand(1)."""
    out = np.zeros(n, dtype=np.int32)
    lib().orc_crand(C.c_int(n), ip(out))
    return out


def params_from(src):
    """An oracle Params holding the fields of an engine cmd_params (a deck
    parsed by the engine) that the oracle has."""
    p = Params()
    names = {f[0] for f in Params._fields_}
    for f in src._fields_:
        if f[0] in names:
            C.memmove(C.addressof(p) + getattr(Params, f[0]).offset,
                      C.addressof(src) + getattr(type(src), f[0]).offset,
                      C.sizeof(f[1]))
    return p


def rng_reset(ibase, coord):
    c = np.ascontiguousarray(coord, dtype=np.float64)
    return lib().orc_rng_reset(C.c_int(ibase), dp(c))


def rng_uniform(seed):
    s = C.c_int(seed)
    u = lib().orc_rng_uniform(C.byref(s))
    return u, s.value


def rng_gaussian(seed):
    s = C.c_int(seed)
    g = lib().orc_rng_gaussian(C.byref(s))
    return g, s.value


def lattice(p, d):
    n = lib().orc_lattice_count(C.byref(p), C.byref(d))
    x = np.zeros((n, 3))
    t = np.zeros(n, dtype=np.int32)
    i = np.zeros(n, dtype=np.int32)
    lib().orc_lattice_fill(C.byref(p), C.byref(d), dp(x), ip(t), ip(i))
    return x, t, i


def velocity_create(p, x, type_):
    n = x.shape[0]
    v = np.zeros((n, 3))
    lib().orc_velocity_create(C.byref(p), C.c_int(n), dp(x), ip(type_), dp(v))
    return v


def nve_initial(x, v, f, type_, mass, dt, mvv2e=1.0):
    n = x.shape[0]
    m = np.ascontiguousarray(mass, dtype=np.float64)
    lib().orc_nve_initial(C.c_int(n), dp(x), dp(v), dp(f), ip(type_), dp(m),
                          C.c_double(dt), C.c_double(mvv2e))


def nve_final(v, f, type_, mass, dt, mvv2e=1.0):
    n = v.shape[0]
    m = np.ascontiguousarray(mass, dtype=np.float64)
    lib().orc_nve_final(C.c_int(n), dp(v), dp(f), ip(type_), dp(m),
                        C.c_double(dt), C.c_double(mvv2e))


def wrap(x, L):
    lib().orc_wrap(C.c_int(x.shape[0]), dp(x), _d3(L))


def bin_sort(x, d, delta, halo_depth=1):
    n = x.shape[0]
    perm = np.zeros(n, dtype=np.int32)
    cell = np.zeros(n, dtype=np.int32)
    nc = lib().orc_bin_sort(C.c_int(n), dp(x), C.byref(d), _d3(delta),
                            C.c_int(halo_depth), ip(perm), ip(cell))
    return perm, cell, nc


def bin_grid(d, delta, halo_depth=1):
    gmin = np.zeros(3)
    rd = np.zeros(3)
    ng = np.zeros(3, dtype=np.int32)
    lib().orc_bin_grid(C.byref(d), _d3(delta), C.c_int(halo_depth), dp(gmin),
                       dp(rd), ip(ng))
    return gmin, rd, ng


def neigh_build_grid(x, n_local, rc, half, gmin, rd, ng, floor_mode=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    gmin = np.ascontiguousarray(gmin, dtype=np.float64)
    rd = np.ascontiguousarray(rd, dtype=np.float64)
    ng = np.ascontiguousarray(ng, dtype=np.int32)
    nt = x.shape[0]
    off = np.zeros(n_local + 1, dtype=np.int64)
    f = lib().orc_neigh_build_grid
    nnz = f(C.c_int(n_local), C.c_int(nt), dp(x), C.c_double(rc),
            C.c_int(int(half)), dp(gmin), dp(rd), ip(ng), C.c_int(floor_mode),
            lp(off), None)
    idx = np.zeros(max(nnz, 1), dtype=np.int32)
    f(C.c_int(n_local), C.c_int(nt), dp(x), C.c_double(rc), C.c_int(int(half)),
      dp(gmin), dp(rd), ip(ng), C.c_int(floor_mode), lp(off), ip(idx))
    return off, idx[:nnz]


def ghost_build(x_local, type_local, id_local, d, depth, cap=None):
    n = x_local.shape[0]
    cap = cap or (4 * n + 4096)
    x = np.zeros((cap, 3))
    t = np.zeros(cap, dtype=np.int32)
    i = np.zeros(cap, dtype=np.int32)
    x[:n] = x_local
    t[:n] = type_local
    i[:n] = id_local
    send = np.zeros(cap, dtype=np.int32)
    counts = np.zeros(6, dtype=np.int32)
    ng = lib().orc_ghost_build(C.c_int(n), C.c_int(cap), dp(x), ip(t), ip(i),
                               C.byref(d), C.c_double(depth), ip(send),
                               ip(counts))
    if ng < 0:
        raise RuntimeError("ghost capacity too small")
    tot = n + ng
    return (x[:tot].copy(), t[:tot].copy(), i[:tot].copy(),
            send[:counts.sum()].copy(), counts)


def ghost_refresh(x, n_local, d, send, counts):
    lib().orc_ghost_refresh(C.c_int(n_local), dp(x), C.byref(d), ip(send),
                            ip(counts))


def ghost_fold(f, n_local, send, counts):
    lib().orc_ghost_fold(C.c_int(n_local), dp(f), ip(send), ip(counts))


def _neigh(fn, x, n_local, rc, half):
    x = np.ascontiguousarray(x, dtype=np.float64)
    nt = x.shape[0]
    off = np.zeros(n_local + 1, dtype=np.int64)
    nnz = fn(C.c_int(n_local), C.c_int(nt), dp(x), C.c_double(rc),
             C.c_int(int(half)), lp(off), None)
    idx = np.zeros(max(nnz, 1), dtype=np.int32)
    fn(C.c_int(n_local), C.c_int(nt), dp(x), C.c_double(rc),
       C.c_int(int(half)), lp(off), ip(idx))
    return off, idx[:nnz]


def neigh_build(x, n_local, rc, half=False):
    return _neigh(lib().orc_neigh_build, x, n_local, rc, half)


def neigh_brute(x, n_local, rc, half=False):
    return _neigh(lib().orc_neigh_brute, x, n_local, rc, half)


def lj_coeff(ntypes, coeffs):
    """coeffs: list of (i, j, eps, sigma, cut) with 1-based types."""
    n = len(coeffs)
    ci = np.array([c[0] for c in coeffs], dtype=np.int32)
    cj = np.array([c[1] for c in coeffs], dtype=np.int32)
    e = np.array([c[2] for c in coeffs], dtype=np.float64)
    s = np.array([c[3] for c in coeffs], dtype=np.float64)
    r = np.array([c[4] for c in coeffs], dtype=np.float64)
    lj1 = np.zeros(ntypes * ntypes)
    lj2 = np.zeros(ntypes * ntypes)
    cs = np.zeros(ntypes * ntypes)
    lib().orc_lj_coeff(C.c_int(ntypes), C.c_int(n), ip(ci), ip(cj), dp(e),
                       dp(s), dp(r), dp(lj1), dp(lj2), dp(cs))
    return lj1, lj2, cs


def lj_force(x, type_, n_local, ntypes, lj1, lj2, cutsq, off, idx, half,
             f=None):
    nt = x.shape[0]
    if f is None:
        f = np.zeros((nt, 3))
    lib().orc_lj_force(C.c_int(n_local), C.c_int(nt), dp(x), ip(type_),
                       C.c_int(ntypes), dp(lj1), dp(lj2), dp(cutsq), lp(off),
                       ip(idx), C.c_int(int(half)), dp(f))
    return f


def lj_energy(x, type_, n_local, ntypes, lj1, lj2, cutsq, off, idx, half):
    return lib().orc_lj_energy(C.c_int(n_local), dp(x), ip(type_),
                               C.c_int(ntypes), dp(lj1), dp(lj2), dp(cutsq),
                               lp(off), ip(idx), C.c_int(int(half)))


def lj_virial(x, type_, n_local, ntypes, lj1, lj2, cutsq, off, idx, half):
    return lib().orc_lj_virial(C.c_int(n_local), dp(x), ip(type_),
                               C.c_int(ntypes), dp(lj1), dp(lj2), dp(cutsq),
                               lp(off), ip(idx), C.c_int(int(half)))


def temperature(v, type_, mass, n_global=None, mvv2e=1.0, boltz=1.0):
    n = v.shape[0]
    m = np.ascontiguousarray(mass, dtype=np.float64)
    return lib().orc_temperature(C.c_int(n), C.c_int(n_global or n), dp(v),
                                 ip(type_), dp(m), C.c_double(mvv2e),
                                 C.c_double(boltz))


def kinetic_energy(v, type_, mass, mvv2e=1.0):
    m = np.ascontiguousarray(mass, dtype=np.float64)
    return lib().orc_kinetic_energy(C.c_int(v.shape[0]), dp(v), ip(type_),
                                    dp(m), C.c_double(mvv2e))


class MD:
    """The reference's CbnMD::init/run loop on the CPU (FP64)."""

    def __init__(self, params):
        self.p = params
        self.h = lib().orc_md_create(C.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_md_destroy(self.h)
            self.h = None

    def init(self):
        lib().orc_md_init(self.h)

    def run(self, nsteps):
        lib().orc_md_run(self.h, nsteps)

    @property
    def n_local(self):
        return lib().orc_md_n_local(self.h)

    @property
    def n_ghost(self):
        return lib().orc_md_n_ghost(self.h)

    @property
    def nnz(self):
        return lib().orc_md_nnz(self.h)

    def arrays(self):
        n = self.n_local + self.n_ghost
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        f = np.zeros((n, 3))
        t = np.zeros(n, dtype=np.int32)
        i = np.zeros(n, dtype=np.int32)
        lib().orc_md_get(self.h, dp(x), dp(v), dp(f), ip(t), ip(i))
        return x, v, f, t, i

    def thermo(self):
        n = lib().orc_md_thermo(self.h, 0, None)
        rows = np.zeros((max(n, 1), 4))
        lib().orc_md_thermo(self.h, n, dp(rows))
        return rows[:n]

    def timers(self):
        t = np.zeros(6)
        lib().orc_md_timers(self.h, dp(t))
        return dict(zip(["force", "neigh", "comm", "int", "other", "total"],
                        t))


# ---- NNP (oracle/nnp.cpp) -----------------------------------------------
class NNP:
    """FP64 Behler-Parrinello model of the oracle (nnp.cpp)."""

    def __init__(self, nelem, hidden, activations, scale=True, smin=0.0,
                 smax=1.0):
        L = lib()
        L.orc_nnp_create.restype = C.c_void_p
        L.orc_nnp_compute.restype = C.c_double
        L.orc_nnp_reference_energy.restype = C.c_double
        hid = (C.c_int * max(len(hidden), 1))(*hidden)
        self.h = C.c_void_p(L.orc_nnp_create(
            C.c_int(nelem), C.c_int(len(hidden) + 1), hid,
            activations.encode()))
        self.nelem = nelem
        self.hidden = list(hidden)
        L.orc_nnp_set_scaling_range(self.h, C.c_int(int(scale)),
                                    C.c_double(smin), C.c_double(smax))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_nnp_destroy(self.h)
            self.h = None

    def add_sf(self, elem, typ, e1, e2, eta, rs, lam, zeta, rc, gmin=0.0,
               gmax=1.0):
        return lib().orc_nnp_add_sf(
            self.h, C.c_int(elem), C.c_int(typ), C.c_int(e1), C.c_int(e2),
            C.c_double(eta), C.c_double(rs), C.c_double(lam),
            C.c_double(zeta), C.c_double(rc), C.c_double(gmin),
            C.c_double(gmax))

    def set_layer(self, elem, layer, W, b):
        W = np.ascontiguousarray(W, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        lib().orc_nnp_set_layer(self.h, C.c_int(elem), C.c_int(layer),
                                C.c_int(W.shape[0]), C.c_int(W.shape[1]),
                                dp(W), dp(b))

    def set_atom_energy(self, elem, e):
        lib().orc_nnp_set_atom_energy(self.h, C.c_int(elem), C.c_double(e))

    def nsf(self, elem):
        return lib().orc_nnp_nsf(self.h, C.c_int(elem))

    def symmetry_functions(self, x, type_, n_local, off, idx, ld=None):
        ld = ld or max(self.nsf(e) for e in range(self.nelem))
        G = np.zeros((n_local, ld))
        lib().orc_nnp_symmetry_functions(
            self.h, C.c_int(n_local), dp(np.ascontiguousarray(x)),
            ip(np.ascontiguousarray(type_, dtype=np.int32)), lp(off),
            ip(idx), C.c_int(ld), dp(G))
        return G

    def compute(self, x, type_, n_local, off, idx, ld=None):
        """(sum E_i, E_i, G, dE/dG, f over all atoms incl. ghosts)."""
        ld = ld or max(self.nsf(e) for e in range(self.nelem))
        nt = x.shape[0]
        G = np.zeros((n_local, ld))
        dG = np.zeros((n_local, ld))
        Ei = np.zeros(n_local)
        f = np.zeros((nt, 3))
        e = lib().orc_nnp_compute(
            self.h, C.c_int(n_local), C.c_int(nt),
            dp(np.ascontiguousarray(x)),
            ip(np.ascontiguousarray(type_, dtype=np.int32)), lp(off),
            ip(idx), C.c_int(ld), dp(G), dp(dG), dp(Ei), dp(f))
        return e, Ei, G, dG, f

    def reference_energy(self, n_local, sum_ei):
        return lib().orc_nnp_reference_energy(self.h, C.c_int(n_local),
                                              C.c_double(sum_ei))
