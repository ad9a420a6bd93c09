"""GPU parity of the LJ timestep path against the FP64 oracle.

Every case runs the HIP engine through the C ABI and the oracle on the same
seeded inputs.  Integer/index results (permutations, ghost lists, neighbour
lists) must match exactly; FP64 results must match exactly where the
operation order is the same (integrator, ghost positions, G=1 forces) and
to the stated relative tolerance otherwise.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _liquid_state(ncells, steps=30, seed=87287, half=False):
    """An fcc lattice melted for a few oracle steps (positions off-lattice)."""
    md = O.MD(O.lj_params(ncells=ncells, seed=seed, half=half, thermo=0))
    md.init()
    md.run(steps)
    x, v, f, t, i = md.arrays()
    n = md.n_local
    return md, x[:n].copy(), v[:n].copy(), t[:n].copy(), i[:n].copy()


def _domain_for(ncells, rho=0.8442):
    a = (4.0 / rho) ** (1.0 / 3.0)
    hi = [a * c for c in ncells]
    p = O.lj_params(ncells=ncells)
    return hi, p.comm_ghost_cutoff


def test_integrator_bit_exact(gpu):
    M = gpu
    rng = np.random.default_rng(342343901)
    n, ng = 1000, 200
    lo, hi = -5.3 * 2.32, 4.7 * 2.32
    x = rng.uniform(lo, hi, (n, 3))
    v = rng.uniform(-1, 1, (n, 3))
    f = rng.uniform(-1, 1, (n, 3))
    s = M.System(1)
    s.set_units(0.005, 1.0, 1.0)
    s.set_mass([1.0])
    s.upload(x, v, f, n_ghost=ng)
    integ = M.Integrator(s)
    integ.initial_integrate()
    integ.final_integrate()
    xg, vg, _, _, _ = s.download()
    xo, vo = x.copy(), v.copy()
    t = np.zeros(n, dtype=np.int32)
    O.nve_initial(xo[:n - ng], vo[:n - ng], f[:n - ng], t, [1.0], 0.005)
    O.nve_final(vo[:n - ng], f[:n - ng], t, [1.0], 0.005)
    assert np.array_equal(xg, xo)
    assert np.array_equal(vg, vo)


def test_integrator_reversibility(gpu):
    """tstIntegrator.hpp:83-137: 100 steps, negate v, 100 steps."""
    M = gpu
    rng = np.random.default_rng(342343901)
    n, ng = 1000, 200
    x = rng.uniform(-5.3 * 2.32, 4.7 * 2.32, (n, 3))
    v = rng.uniform(-1, 1, (n, 3))
    f = rng.uniform(-1, 1, (n, 3))
    s = M.System(1)
    s.set_units(0.005, 1.0, 1.0)
    s.upload(x, v, f, n_ghost=ng)
    integ = M.Integrator(s)
    for _ in range(100):
        integ.initial_integrate()
        integ.final_integrate()
    xs, vs, fs, _, _ = s.download()
    s.upload(xs, -vs, fs, n_ghost=ng)
    for _ in range(100):
        integ.initial_integrate()
        integ.final_integrate()
    xf = s.download()[0]
    # EXPECT_FLOAT_EQ: within 4 ULPs of float
    tol = 4 * np.finfo(np.float32).eps * np.maximum(np.abs(x), 1.0)
    assert np.all(np.abs(xf - x) <= tol)


@pytest.mark.parametrize("ncells", [(6, 6, 6), (10, 9, 8)])
def test_binning_permutation_exact(gpu, ncells):
    M = gpu
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    rng = np.random.default_rng(7)
    order = rng.permutation(len(x))  # scramble so the sort has work to do
    x, v, t, i = x[order], v[order], t[order], i[order]
    d = O.create_domain([0, 0, 0], hi, gc)
    O.wrap(x, d.L)
    perm_o, cell_o, _ = O.bin_sort(x, d, [2.8] * 3, 1)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    perm = M.Binning(s).create_binning(2.8, 2.8, 2.8, 1, True, False, True)
    assert np.array_equal(perm, perm_o)
    xs, vs, _, _, ids = s.download()
    assert np.array_equal(xs, x[perm_o])
    assert np.array_equal(vs, v[perm_o])
    assert np.array_equal(ids, i[perm_o])
    assert np.all(np.diff(cell_o) >= 0)


@pytest.mark.parametrize("ncells", [(5, 5, 5), (8, 7, 6)])
def test_ghosts_exact(gpu, ncells):
    M = gpu
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    d = O.create_domain([0, 0, 0], hi, gc)
    O.wrap(x, d.L)
    xo, to, io, send_o, counts_o = O.ghost_build(x, t, i, d, 2.8)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    comm = M.Comm(s, 2.8)
    comm.exchange()
    comm.exchange_halo()
    send, counts = comm.ghost_lists()
    assert np.array_equal(counts, counts_o)
    assert np.array_equal(send, send_o)
    xg, _, _, tg, ig = s.download()
    assert np.array_equal(xg, xo)
    assert np.array_equal(ig, io)
    # refresh after a drift of the local atoms
    s.set_units(0.005)
    integ = M.Integrator(s)
    integ.initial_integrate()
    comm.update_halo()
    xg2 = s.download()[0]
    xl = xo[:len(x)].copy()
    vl = v.copy()
    O.nve_initial(xl, vl, np.zeros_like(xl), t, [1.0], 0.005)
    xo2 = xo.copy()
    xo2[:len(x)] = xl
    O.ghost_refresh(xo2, len(x), d, send_o, counts_o)
    assert np.array_equal(xg2, xo2)
    # the ghost force fold is checked with real half-list forces in
    # test_lj_half_fp64


def _tstneighbor_positions(seed=342343901):
    rc = 2.32
    rng = np.random.default_rng(seed)
    x = rng.uniform(-5.3 * rc, 4.7 * rc, (1000, 3))
    return x, rc


@pytest.mark.parametrize("half", [False, True])
def test_neighbor_tstneighbor_geometry(gpu, half):
    """tstNeighbor.hpp:257-320: 1000 atoms, last 200 ghosts, rc 2.32."""
    M = gpu
    x, rc = _tstneighbor_positions()
    s = M.System(1)
    s.upload(x, n_ghost=200)
    nl = M.NeighborVerlet(rc, half, 100)
    nl.create(s)
    off, idx = nl.get()
    off_o, idx_o = O.neigh_build(x, 800, rc, half)
    assert np.array_equal(off, off_o)
    assert np.array_equal(idx, idx_o)
    off_b, idx_b = O.neigh_brute(x, 800, rc, half)
    for r in range(800):
        a = np.sort(idx[off[r]:off[r + 1]])
        b = np.sort(idx_b[off_b[r]:off_b[r + 1]])
        assert np.array_equal(a, b)
    if half:
        offf, idxf = O.neigh_brute(x, 800, rc, False)
        assert len(idx) <= len(idxf)
        # no pair stored in both directions among local rows
        pairs = set()
        for r in range(800):
            for j in idx[off[r]:off[r + 1]]:
                assert (int(j), r) not in pairs
                pairs.add((r, int(j)))


@pytest.mark.parametrize("half", [False, True])
def test_neighbor_lattice_exact(gpu, half):
    M = gpu
    ncells = (9, 8, 7)
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    d = O.create_domain([0, 0, 0], hi, gc)
    O.wrap(x, d.L)
    xo, _, _, _, _ = O.ghost_build(x, t, i, d, 2.8)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    comm = M.Comm(s, 2.8)
    comm.exchange_halo()
    nl = M.NeighborVerlet(2.8, half, 50)
    nl.create(s)
    off, idx = nl.get()
    off_o, idx_o = O.neigh_build(xo, len(x), 2.8, half)
    assert np.array_equal(off, off_o)
    assert np.array_equal(idx, idx_o)
    assert nl.max_neigh_guess >= int(np.diff(off).max())


@pytest.mark.parametrize("width", [1, 2, 3])
@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
def test_neighbor_on_binning_grid_exact(gpu, half, lanes, width):
    """After create_binning the list is searched on the binning grid; rows
    must equal the oracle's build on that grid, in any slab layout."""
    M = gpu
    ncells = (9, 8, 7)
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    d = O.create_domain([0, 0, 0], hi, gc)
    O.wrap(x, d.L)
    perm, _, _ = O.bin_sort(x, d, [2.8] * 3, 1)
    xo = x[perm]
    xg, _, _, _, _ = O.ghost_build(xo, t[perm], i[perm], d, 2.8)
    gmin, rd, ng = O.bin_grid(d, [2.8] * 3, 1)
    off_o, idx_o = O.neigh_build_grid(xg, len(x), 2.8, half, gmin, rd, ng, 1)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    M.Binning(s).create_binning(2.8, 2.8, 2.8, 1, True, False, True)
    comm = M.Comm(s, 2.8)
    comm.exchange_halo()
    nl = M.NeighborVerlet(2.8, half, 50, lanes=lanes, tile_width=width)
    nl.create(s)
    off, idx = nl.get()
    lay = nl.layout()
    # a half list on 3 x 3 tiles stages FP64 records and reactions for 25
    # columns, which may exceed a block's LDS at every height (then the
    # global layout): only its rows are checked
    if not (half and width == 3):
        assert lay["tiled"] == (lanes > 1)
        assert lay["tile_width"] == (width if lanes > 1 else 0)
    assert np.array_equal(s.download()[0], xg)
    assert np.array_equal(off, off_o)
    assert np.array_equal(idx, idx_o)


@pytest.mark.parametrize("half,width", [(False, 1), (False, 2), (False, 3), (True, 1),
                                        (True, 2)])
def test_list_build_band_retest_exact(gpu, monkeypatch, half, width):
    """The tiled build tests candidates in FP32 and sends those whose FP32
    r^2 lies within a band of rc^2 (and, for half lists, those whose x
    direction the FP32 difference cannot settle) to an FP64 re-test on the
    positions in memory.  With the band widened to 2 (most candidates
    between r = 2.43 and 3.13 re-tested) the rows still equal the oracle's."""
    monkeypatch.setenv("CMD_BUILD_BAND", "2.0")
    M = gpu
    ncells = (9, 8, 7)
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    d = O.create_domain([0, 0, 0], hi, gc)
    O.wrap(x, d.L)
    perm, _, _ = O.bin_sort(x, d, [2.8] * 3, 1)
    xo = x[perm]
    xg, _, _, _, _ = O.ghost_build(xo, t[perm], i[perm], d, 2.8)
    gmin, rd, ng = O.bin_grid(d, [2.8] * 3, 1)
    off_o, idx_o = O.neigh_build_grid(xg, len(x), 2.8, half, gmin, rd, ng, 1)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    M.Binning(s).create_binning(2.8, 2.8, 2.8, 1, True, False, True)
    M.Comm(s, 2.8).exchange_halo()
    nl = M.NeighborVerlet(2.8, half, 50, lanes=2, tile_width=width)
    nl.create(s)
    off, idx = nl.get()
    assert nl.layout()["tiled"]
    assert np.array_equal(off, off_o)
    assert np.array_equal(idx, idx_o)


@pytest.mark.parametrize("shape", ["1x0", "1x8", "2x0", "2x2", "3x0", "3x2"])
@pytest.mark.parametrize("wrapped", [True, False])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lj_tiled_force(gpu, prec, lanes, wrapped, shape):
    """Full list built on the binning grid with 2+ lanes per atom uses the
    LDS-tiled layout when every local atom lies in an interior cell; atoms
    left unwrapped in the padding cells make the build use the global layout.
    Forces and energy against the oracle either way."""
    M = gpu
    ncells = (10, 9, 8)
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    if wrapped:
        O.wrap(x, O.create_domain([0, 0, 0], hi, gc).L)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    M.Binning(s).create_binning(2.8, 2.8, 2.8, 1, True, False, True)
    comm = M.Comm(s, 2.8)
    comm.exchange_halo()
    width, tile = (int(v) for v in shape.split("x"))
    nl = M.NeighborVerlet(2.8, False, 50, lanes=lanes, tile=tile, tile_width=width)
    nl.create(s)
    lay = nl.layout()
    assert lay["tiled"] == wrapped
    if wrapped:
        assert lay["tile_width"] == width
        want = tile or (4 if width == 1 else 3)
        if width == 3:
            # 25 staged columns of FP64 records: the height is lowered until
            # the tile fits a block's LDS
            assert 1 <= lay["tile_height"] <= want
        else:
            assert lay["tile_height"] == want
    F = M.ForceLJ(s, precision=M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED)
    F.init_coeff([["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]])
    F.compute(s, nl)
    xg, _, fg, tg, _ = s.download()
    off, idx = nl.get()
    n = len(x)
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(xg, tg, n, 1, lj1, lj2, cs, off, idx, False)
    err = np.linalg.norm(fg[:n] - fo[:n]) / np.linalg.norm(fo[:n])
    assert err < (1e-13 if prec == "fp64" else 1e-5), err
    eo = O.lj_energy(xg, tg, n, 1, lj1, lj2, cs, off, idx, False)
    assert abs(F.compute_energy(s, nl) - eo) <= 1e-12 * abs(eo)
    # ghosts refreshed through the sorted copy stay consistent
    integ = M.Integrator(s)
    s.set_units(0.005)
    integ.initial_integrate()
    comm.update_halo()
    F.compute(s, nl)
    xg2, _, fg2, _, _ = s.download()
    fo2 = O.lj_force(xg2, tg, n, 1, lj1, lj2, cs, off, idx, False)
    err2 = np.linalg.norm(fg2[:n] - fo2[:n]) / np.linalg.norm(fo2[:n])
    assert err2 < (1e-13 if prec == "fp64" else 1e-5), err2


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lj_tiled_half_force(gpu, prec, lanes):
    """Half list on the binning grid with 2+ lanes per atom: the tiled
    kernel sums reactions per staged atom in LDS and adds them to f with
    atomics; after the ghost fold, forces, energy and virial against the
    oracle's half-list force and fold (force_lj_cabana_neigh_impl.h:205-259,
    comm_mpi_impl.h:410-441)."""
    M = gpu
    ncells = (10, 9, 8)
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    O.wrap(x, O.create_domain([0, 0, 0], hi, gc).L)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    M.Binning(s).create_binning(2.8, 2.8, 2.8, 1, True, False, True)
    comm = M.Comm(s, 2.8)
    comm.exchange_halo()
    nl = M.NeighborVerlet(2.8, True, 50, lanes=lanes)
    nl.create(s)
    F = M.ForceLJ(s, precision=M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED)
    F.init_coeff([["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]])
    F.compute(s, nl)
    comm.update_force()
    xg, _, fg, tg, _ = s.download()
    off, idx = nl.get()
    n = len(x)
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(xg, tg, n, 1, lj1, lj2, cs, off, idx, True)
    send, counts = comm.ghost_lists()
    O.ghost_fold(fo, n, send, counts)
    err = np.linalg.norm(fg[:n] - fo[:n]) / np.linalg.norm(fo[:n])
    assert err < (1e-12 if prec == "fp64" else 1e-5), err
    assert np.abs(fg[:n].sum(0)).max() < 1e-8
    eo = O.lj_energy(xg, tg, n, 1, lj1, lj2, cs, off, idx, True)
    wo = O.lj_virial(xg, tg, n, 1, lj1, lj2, cs, off, idx, True)
    assert abs(F.compute_energy(s, nl) - eo) <= 1e-12 * abs(eo)
    assert abs(F.compute_virial(s, nl) - wo) <= 1e-12 * abs(wo)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_loop_half_tiled(gpu, prec):
    """CbnMD with a half list and two lanes per atom (the tiled half kernel,
    its thermo energy with the reference's 1/2 weight for ghost pairs, and
    the ghost fold) against the FP64 oracle's loop."""
    M = gpu
    kw = dict(ncells=(10, 10, 10), nsteps=100, thermo=10, half=True)
    md = M.CbnMD(M.lattice_params(**kw),
                 precision=M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED,
                 lanes=2)
    md.init()
    md.run(100)
    ref = O.MD(O.lj_params(**kw))
    ref.init()
    ref.run(100)
    th, tho = md.thermo(), ref.thermo()
    assert th.shape[0] == tho.shape[0] == 11
    x, v, f, _, _ = md.arrays()
    xo, vo, fo, _, _ = ref.arrays()
    n = md.n_local
    if prec == "fp64":
        assert np.allclose(th[:, :4], tho, rtol=1e-10, atol=1e-13)
        assert np.abs(x[:n] - xo[:n]).max() < 1e-9
        assert np.abs(v[:n] - vo[:n]).max() < 1e-9
    else:
        assert np.allclose(th[:, :4], tho, rtol=1e-5, atol=1e-7)
        assert np.abs(x[:n] - xo[:n]).max() < 1e-5


def _force_case(M, ncells, half, prec, lanes):
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    d = O.create_domain([0, 0, 0], hi, gc)
    O.wrap(x, d.L)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    comm = M.Comm(s, 2.8)
    comm.exchange_halo()
    nl = M.NeighborVerlet(2.8, half, 50, lanes=lanes)
    nl.create(s)
    F = M.ForceLJ(s, precision=prec)
    F.init_coeff([["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]])
    F.compute(s, nl)
    if half:
        comm.update_force()
    xg, _, fg, tg, _ = s.download()
    off, idx = nl.get()
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(xg, tg, len(x), 1, lj1, lj2, cs, off, idx, half)
    send, counts = comm.ghost_lists()
    if half:
        O.ghost_fold(fo, len(x), send, counts)
    eo = O.lj_energy(xg, tg, len(x), 1, lj1, lj2, cs, off, idx, half)
    wo = O.lj_virial(xg, tg, len(x), 1, lj1, lj2, cs, off, idx, half)
    e = F.compute_energy(s, nl)
    w = F.compute_virial(s, nl)
    return fg[:len(x)], fo[:len(x)], e, eo, w, wo


def test_lj_full_fp64_bit_exact_one_lane(gpu):
    fg, fo, e, eo, w, wo = _force_case(gpu, (8, 8, 8), False, gpu.PREC_FP64, 1)
    assert np.array_equal(fg, fo)
    assert abs(e - eo) <= 1e-12 * abs(eo)
    assert abs(w - wo) <= 1e-12 * abs(wo)


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_lj_full_fp64_lanes(gpu, lanes):
    fg, fo, e, eo, w, wo = _force_case(gpu, (8, 8, 8), False, gpu.PREC_FP64,
                                       lanes)
    err = np.linalg.norm(fg - fo) / np.linalg.norm(fo)
    assert err < 1e-13
    assert np.abs(fg.sum(0)).max() < 1e-9


def test_lj_half_fp64(gpu):
    fg, fo, e, eo, w, wo = _force_case(gpu, (8, 8, 8), True, gpu.PREC_FP64, 1)
    err = np.linalg.norm(fg - fo) / np.linalg.norm(fo)
    assert err < 1e-12
    assert abs(e - eo) <= 1e-12 * abs(eo)
    assert abs(w - wo) <= 1e-12 * abs(wo)


def test_lj_mixed_precision(gpu):
    fg, fo, e, eo, w, wo = _force_case(gpu, (8, 8, 8), False, gpu.PREC_MIXED,
                                       1)
    err = np.linalg.norm(fg - fo) / np.linalg.norm(fo)
    rms = np.sqrt((fo ** 2).sum(1).mean())
    maxe = np.abs(fg - fo).max() / rms
    print(f"mixed: |dF|/|F| = {err:.3e}, max|dF_i|/rms|F| = {maxe:.3e}")
    assert err < 1e-5
    assert maxe < 1e-4


@pytest.mark.parametrize("half", [False, True])
def test_md_loop_fp64(gpu, half):
    """CbnMD::init + run on fcc 10^3 cells (4000 atoms), 100 steps."""
    M = gpu
    kw = dict(ncells=(10, 10, 10), nsteps=100, thermo=10, half=half)
    md = M.CbnMD(M.lattice_params(**kw), precision=M.PREC_FP64, lanes=1)
    md.init()
    md.run(100)
    ref = O.MD(O.lj_params(**kw))
    ref.init()
    ref.run(100)
    th, tho = md.thermo(), ref.thermo()
    assert th.shape[0] == tho.shape[0] == 11
    assert np.allclose(th[:, :4], tho, rtol=1e-10, atol=1e-13)
    x, v, f, _, _ = md.arrays()
    xo, vo, fo, _, _ = ref.arrays()
    n = md.n_local
    if not half:
        # same operations in the same order: bit-identical trajectory
        assert np.array_equal(x[:n], xo[:n])
        assert np.array_equal(v[:n], vo[:n])
    else:
        assert np.abs(x[:n] - xo[:n]).max() < 1e-9
        assert np.abs(v[:n] - vo[:n]).max() < 1e-9


@pytest.mark.parametrize("lanes", [2, 4])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_md_fused_ghost_refresh_bit_identical(gpu, prec, lanes):
    """On tiled full lists the fused step writes the next step's sorted
    ghosts (update_halo, comm_mpi_impl.h:369-408, inside the step): over
    rebuilds, thermo steps and a split run, x, v, f (ghost rows included)
    and thermo rows equal those of the separate refresh pass bit for bit,
    with one lane finishing an atom (2 lanes) and one lane per component
    (4 lanes)."""
    M = gpu
    kw = dict(ncells=(10, 9, 8), nsteps=150, thermo=10)
    P = M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED
    a = M.CbnMD(M.lattice_params(**kw), precision=P, lanes=lanes, fuse_ghosts=True)
    b = M.CbnMD(M.lattice_params(**kw), precision=P, lanes=lanes, fuse_ghosts=False)
    a.init()
    b.init()
    for k in (45, 105):
        a.run(k)
        b.run(k)
        for u, w in zip(a.arrays(), b.arrays()):
            assert np.array_equal(u, w)
    assert np.array_equal(a.thermo(), b.thermo())


def test_md_loop_split_runs_match_one_run(gpu):
    M = gpu
    kw = dict(ncells=(6, 6, 6), nsteps=50, thermo=10)
    a = M.CbnMD(M.lattice_params(**kw), lanes=1)
    a.init()
    a.run(50)
    b = M.CbnMD(M.lattice_params(**kw), lanes=1)
    b.init()
    b.run(17)
    b.run(33)
    xa, va = a.arrays()[:2]
    xb, vb = b.arrays()[:2]
    n = a.n_local
    assert np.array_equal(xa[:n], xb[:n])
    assert np.array_equal(va[:n], vb[:n])


def test_md_mixed_energy_drift(gpu):
    M = gpu
    kw = dict(ncells=(10, 10, 10), nsteps=200, thermo=20)
    md = M.CbnMD(M.lattice_params(**kw), precision=M.PREC_MIXED)
    md.init()
    md.run(200)
    th = md.thermo()
    e = th[:, 3]
    drift = np.abs(e - e[0]).max() / abs(e[0])
    print(f"mixed precision drift over 200 steps: {drift:.3e}")
    assert drift < 1e-3


MULTI = [["pair_coeff", "1", "1", "1.0", "1.0", "2.5"],
         ["pair_coeff", "1", "2", "0.8", "1.1", "2.4"],
         ["pair_coeff", "2", "2", "1.3", "0.9", "2.2"]]


@pytest.mark.parametrize("path", ["global1", "tiled", "tiled-e8", "half"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_lj_two_types(gpu, prec, path):
    """Two atom types with their own eps, sigma and cutoff per pair
    (force_lj_cabana_neigh_impl.h:62-89 symmetric tables): forces and energy
    against the oracle on the global and the tiled layouts."""
    M = gpu
    ncells = (9, 9, 8)
    _, x, v, t, i = _liquid_state(ncells)
    t = (i % 2).astype(np.int32)  # alternate types by atom id
    hi, gc = _domain_for(ncells)
    O.wrap(x, O.create_domain([0, 0, 0], hi, gc).L)
    s = M.System(2)
    s.create_domain([0, 0, 0], hi, gc)
    s.set_mass([1.0, 1.7])
    s.upload(x, v, None, t, i)
    M.Binning(s).create_binning(2.8, 2.8, 2.8, 1, True, False, True)
    comm = M.Comm(s, 2.8)
    comm.exchange_halo()
    half = path == "half"
    lanes = 1 if path == "global1" else 2
    nl = M.NeighborVerlet(2.8, half, 50, lanes=lanes,
                          tile=(8 if path == "tiled-e8" else 0))
    nl.create(s)
    F = M.ForceLJ(s, precision=M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED)
    F.init_coeff(MULTI)
    F.compute(s, nl)
    if half:
        comm.update_force()
    xg, _, fg, tg, _ = s.download()
    off, idx = nl.get()
    n = len(x)
    lj1, lj2, cs = O.lj_coeff(2, [(int(c[1]), int(c[2]), float(c[3]),
                                   float(c[4]), float(c[5])) for c in MULTI])
    fo = O.lj_force(xg, tg, n, 2, lj1, lj2, cs, off, idx, half)
    if half:
        send, counts = comm.ghost_lists()
        O.ghost_fold(fo, n, send, counts)
    err = np.linalg.norm(fg[:n] - fo[:n]) / np.linalg.norm(fo[:n])
    assert err < (1e-12 if prec == "fp64" else 1e-5), err
    eo = O.lj_energy(xg, tg, n, 2, lj1, lj2, cs, off, idx, half)
    assert abs(F.compute_energy(s, nl) - eo) <= 1e-10 * abs(eo)


def test_ghost_rows_after_halo(gpu):
    """Ghost rows made by the halo exchange: positions are the root's image
    (comm_mpi_impl.h:346-368), types follow the root, velocities read as
    zero (the halo gathers x and type only), both after init and after a
    run whose steps wrote the sorted ghosts inside the fused step."""
    M = gpu
    kw = dict(ncells=(8, 8, 8), nsteps=40, thermo=10)
    md = M.CbnMD(M.lattice_params(**kw), lanes=2)
    md.init()
    for k in (0, 33):
        if k:
            md.run(k)
        x, v, f, t, i = md.arrays()
        n = md.n_local
        assert len(x) > n
        assert np.array_equal(v[n:], np.zeros_like(v[n:]))
        L = 8 * (4 / 0.8442) ** (1 / 3)  # 8 fcc cells at rho 0.8442
        # every ghost is a local atom shifted by a whole number of box lengths
        by_id = {int(a): r for r, a in enumerate(i[:n])}
        src = np.array([by_id[int(a)] for a in i[n:]])
        shift = (x[n:] - x[src]) / L
        assert np.allclose(shift, np.round(shift), atol=1e-9)
        assert np.array_equal(t[n:], t[src])


def test_init_tuning_restores_initial_state(gpu):
    """Init chooses the list layout on a scratch run (20 steps, layouts
    timed at its list rebuild), then uploads the initial state again: x and
    v after init equal an untuned init's, the step-0 thermo row agrees and
    the run starts at step 0."""
    M = gpu
    kw = dict(ncells=(12, 12, 12), nsteps=40, thermo=20)
    a = M.CbnMD(M.lattice_params(**kw), precision=M.PREC_FP64)
    a.init()
    b = M.CbnMD(M.lattice_params(**kw), precision=M.PREC_FP64, lanes=1)
    b.init()
    g, tile, secs, width = a.tune_info()
    print(f"tuned layout: lanes {g}, tile {width}x{width}x{tile}, {secs:.3f} s", flush=True)
    assert g in (1, 2, 4, 8) and secs > 0.0
    xa, va, _, _, ia = a.arrays()
    xb, vb, _, _, ib = b.arrays()
    n = a.n_local
    assert n == b.n_local
    assert np.array_equal(ia[:n], ib[:n])
    assert np.array_equal(xa[:n], xb[:n])
    assert np.array_equal(va[:n], vb[:n])
    ta, tb = a.thermo(), b.thermo()
    assert ta.shape == tb.shape == (1, 5)
    assert np.allclose(ta, tb, rtol=1e-12, atol=1e-14)
    a.run(40)
    b.run(40)
    ta, tb = a.thermo(), b.thermo()
    assert np.array_equal(ta[:, 0], [0, 20, 40])
    assert np.allclose(ta[:, :4], tb[:, :4], rtol=1e-10, atol=1e-13)


def test_halo_zeroes_ghost_velocity_and_force_rows(gpu):
    """Comm::exchange_halo gathers positions and types into the ghosts
    (comm_mpi_impl.h:346-347); the engine sets their velocity and force rows
    to zero, also over force rows a half-list force wrote before."""
    M = gpu
    ncells = (6, 6, 6)
    _, x, v, t, i = _liquid_state(ncells)
    hi, gc = _domain_for(ncells)
    s = M.System(1)
    s.create_domain([0, 0, 0], hi, gc)
    s.upload(x, v, None, t, i)
    M.Binning(s).create_binning(2.8, 2.8, 2.8, 1, True, False, True)
    comm = M.Comm(s, 2.8)
    comm.exchange_halo()
    nl = M.NeighborVerlet(2.8, True, 50, lanes=1)
    nl.create(s)
    F = M.ForceLJ(s, precision=M.PREC_FP64)
    F.init_coeff([["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]])
    F.compute(s, nl)  # a half list on the global slab: reactions on ghost rows
    n = len(x)
    _, _, f1, _, _ = s.download()
    assert len(f1) > n and np.abs(f1[n:]).max() > 0
    comm.exchange_halo()
    _, v2, f2, _, _ = s.download()
    assert len(f2) > n
    assert not np.any(v2[n:]) and not np.any(f2[n:])
