"""GPU edge cases of the LJ path against the FP64 oracle: dense boxes (cells
of more than 64 atoms, stencils of more than 640 candidates), boxes with
empty regions, boxes smaller than twice the halo depth, and empty systems.
Lists and ghosts must match exactly, forces to the stated tolerance."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

RC = 2.8


def _grid_lists(x, L, half):
    """The oracle's build on the binning grid of create_binning, with the
    binned order and the six-phase ghosts."""
    d = O.create_domain([0, 0, 0], L, RC)
    x = x.copy()
    O.wrap(x, d.L)
    perm, _, _ = O.bin_sort(x, d, [RC] * 3, 1)
    n = len(x)
    t = np.zeros(n, np.int32)
    ids = np.arange(1, n + 1, dtype=np.int32)
    xg, _, _, send, counts = O.ghost_build(x[perm], t, ids[perm], d, RC)
    gmin, rd, ng = O.bin_grid(d, [RC] * 3, 1)
    off, idx = O.neigh_build_grid(xg, n, RC, half, gmin, rd, ng, 1)
    return x, xg, off, idx, send, counts


def _engine(M, x, L, half, lanes, guess=50, width=0):
    s = M.System(1)
    s.create_domain([0, 0, 0], L, RC)
    s.upload(x, np.zeros_like(x), None, np.zeros(len(x), np.int32),
             np.arange(1, len(x) + 1, dtype=np.int32))
    M.Binning(s).create_binning(RC, RC, RC, 1, True, False, True)
    comm = M.Comm(s, RC)
    comm.exchange_halo()
    nl = M.NeighborVerlet(RC, half, guess, lanes=lanes, tile_width=width)
    nl.create(s)
    return s, comm, nl


@pytest.mark.parametrize("width", [1, 2])
@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("lanes", [2, 8])
def test_dense_box_lists_exact(gpu, half, lanes, width):
    """Density 2.4 (about 53 atoms per cell, some above 64; about 1,400
    stencil candidates per cell, so the tiled build takes two register
    passes), rows of ~220: lists equal to the oracle's, entry for entry."""
    M = gpu
    rng = np.random.default_rng(20240611)
    L = [8 * RC + 0.05] * 3
    n = int(2.4 * np.prod(L))
    x = rng.uniform(0.0, 1.0, (n, 3)) * np.array(L)
    xw, xg, off_o, idx_o, _, _ = _grid_lists(x, L, half)
    s, comm, nl = _engine(M, x, L, half, lanes, guess=20, width=width)
    lay = nl.layout()
    print("layout", lay, "rows up to", nl.info()["max_row"])
    assert lay["tiled"]  # the register-pass build, not the global fallback
    off, idx = nl.get()
    assert np.array_equal(s.download()[0], xg)
    assert np.array_equal(off, off_o)
    assert np.array_equal(idx, idx_o)


@pytest.mark.parametrize("half", [False, True])
def test_vacuum_slab_forces(gpu, half):
    """A melted liquid with the lower half of the box in x emptied: tiles
    over empty columns and half-empty stencils; FP64 forces after the ghost
    fold against the oracle."""
    M = gpu
    ncells = (10, 9, 8)
    md = O.MD(O.lj_params(ncells=ncells, thermo=0))
    md.init()
    md.run(30)
    x, _, _, _, _ = md.arrays()
    x = x[:md.n_local].copy()
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    L = [a * c for c in ncells]
    O.wrap(x, O.create_domain([0, 0, 0], L, RC).L)
    x = x[x[:, 0] >= 0.5 * L[0]]
    xw, xg, off_o, idx_o, send, counts = _grid_lists(x, L, half)
    s, comm, nl = _engine(M, x, L, half, 2)
    off, idx = nl.get()
    assert np.array_equal(off, off_o)
    assert np.array_equal(idx, idx_o)
    F = M.ForceLJ(s, precision=M.PREC_FP64)
    F.init_coeff([["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]])
    F.compute(s, nl)
    if half:
        comm.update_force()
    xgg, _, fg, tg, _ = s.download()
    n = len(x)
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(xgg, tg, n, 1, lj1, lj2, cs, off, idx, half)
    if half:
        O.ghost_fold(fo, n, send, counts)
    err = np.linalg.norm(fg[:n] - fo[:n]) / np.linalg.norm(fo[:n])
    assert err < 1e-12, err


@pytest.mark.parametrize("lanes", [1, 2])
def test_box_smaller_than_two_halo_depths(gpu, lanes):
    """fcc 2 x 2 x 2 cells (32 atoms, L = 3.36 < 2 x 2.8): atoms near both
    faces of a dimension have images on both sides; ghosts, send lists and
    counts exact, forces and energy against the oracle."""
    M = gpu
    md = O.MD(O.lj_params(ncells=(2, 2, 2), thermo=0))
    md.init()
    md.run(10)
    x, _, _, _, _ = md.arrays()
    n = md.n_local
    x = x[:n].copy()
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    L = [2 * a] * 3
    d = O.create_domain([0, 0, 0], L, RC)
    O.wrap(x, d.L)
    t = np.zeros(n, np.int32)
    ids = np.arange(1, n + 1, dtype=np.int32)
    xo, _, io, send_o, counts_o = O.ghost_build(x, t, ids, d, RC)
    s = M.System(1)
    s.create_domain([0, 0, 0], L, RC)
    s.upload(x, np.zeros_like(x), None, t, ids)
    comm = M.Comm(s, RC)
    comm.exchange_halo()
    send, counts = comm.ghost_lists()
    assert np.array_equal(counts, counts_o)
    assert np.array_equal(send, send_o)
    xg, _, _, _, ig = s.download()
    assert np.array_equal(xg, xo)
    assert np.array_equal(ig, io)
    nl = M.NeighborVerlet(RC, False, 50, lanes=lanes)
    nl.create(s)
    F = M.ForceLJ(s, precision=M.PREC_FP64)
    F.init_coeff([["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]])
    F.compute(s, nl)
    _, _, fg, tg, _ = s.download()
    off, idx = nl.get()
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(xg, tg, n, 1, lj1, lj2, cs, off, idx, False)
    assert np.abs(fg[:n] - fo[:n]).max() <= 1e-12 * np.abs(fo[:n]).max()
    eo = O.lj_energy(xg, tg, n, 1, lj1, lj2, cs, off, idx, False)
    assert abs(F.compute_energy(s, nl) - eo) <= 1e-12 * abs(eo)


def test_empty_system(gpu):
    """No atoms: the halo, list, force and integrator steps succeed and
    leave nothing."""
    M = gpu
    s = M.System(1)
    s.create_domain([0, 0, 0], [10.0, 10.0, 10.0], RC)
    s.set_units(0.005)
    s.upload(np.zeros((0, 3)))
    comm = M.Comm(s, RC)
    comm.exchange()
    comm.exchange_halo()
    send, counts = comm.ghost_lists()
    assert counts.sum() == 0
    nl = M.NeighborVerlet(RC, False, 50, lanes=2)
    nl.create(s)
    off, idx = nl.get()
    assert len(idx) == 0
    F = M.ForceLJ(s, precision=M.PREC_FP64)
    F.init_coeff([["pair_coeff", "1", "1", "1.0", "1.0", "2.5"]])
    F.compute(s, nl)
    M.Integrator(s).initial_integrate()
    assert s.counts() == (0, 0, 0)
