"""Batched ensemble (cmd_batch) on the GPU, through the C ABI.

Each system of a batch follows CbnMD::run (cabanamd_impl.h:285-399), so it
is checked against the FP64 oracle's MD loop started from the same seeded
fcc state: positions and velocities after 50 steps (two rebuilds) within
1e-9 and thermo within 1e-10 relative in FP64 (the batch sums forces in its
own neighbour order, the oracle in the reference's; the difference starts
at rounding level); forces on a liquid state within 1e-5 relative in mixed
precision.
A system's trajectory must not depend on the rest of the batch: alone and
in a batch whose other members join and leave around it, its x, v, f and
thermo rows are bit-identical.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

# (cells, density, temperature, seed): uneven sizes and densities
SYSTEMS = [((5, 5, 5), 0.70, 0.8, 101), ((6, 5, 7), 0.85, 1.6, 102),
           ((7, 7, 6), 0.90, 2.0, 103), ((5, 6, 5), 0.78, 1.2, 104)]


def _params(M, thermo=20):
    return M.lattice_params(ncells=(5, 5, 5), thermo=thermo, every=20)


def _add(M, B, spec):
    cells, rho, temp, seed = spec
    x, v, t, i, lo, hi = M.fcc_system(cells, rho, temp, seed)
    return B.add(x, t, lo, hi, v=v, id=i)


def _oracle(spec, nsteps):
    cells, rho, temp, seed = spec
    md = O.MD(O.lj_params(ncells=cells, rho=rho, temp=temp, seed=seed,
                          gaussian=True, thermo=10, every=20))
    md.init()
    md.run(nsteps)
    x, v, f, t, i = md.arrays()
    n = md.n_local
    o = np.argsort(i[:n], kind="stable")
    return x[:n][o], v[:n][o], f[:n][o], md.thermo()


def _min_image(d, L):
    return d - L * np.round(d / L)


def test_batch_fp64_matches_oracle(gpu):
    M = gpu
    B = M.Batch(_params(M), M.PREC_FP64, lanes=2)
    hs = [_add(M, B, s) for s in SYSTEMS]
    B.run(50)
    assert B.counts()["systems"] == len(SYSTEMS)
    for h, spec in zip(hs, SYSTEMS):
        cells, rho = spec[0], spec[1]
        L = np.array(cells) * (4.0 / rho) ** (1.0 / 3.0)
        x, v, f, _, _ = B.download(h)
        xo, vo, fo, tho = _oracle(spec, 50)
        assert x.shape == xo.shape
        assert np.abs(_min_image(x - xo, L)).max() < 1e-9
        assert np.abs(v - vo).max() < 1e-9
        assert np.abs(f - fo).max() < 1e-7 * np.abs(fo).max()
        th = B.thermo(h)
        assert list(th[:, 0]) == [0, 20, 40, 50]
        for row in th:
            ro = tho[tho[:, 0] == row[0]][0]
            assert np.allclose(row[1:], ro[1:], rtol=1e-10, atol=1e-12), (row, ro)


def test_batch_mixed_forces_within_1e5(gpu):
    """Forces of a mixed-precision batch on the oracle's state after 30
    steps (off the lattice, where forces no longer cancel) against the
    oracle's FP64 forces on the same positions."""
    M = gpu
    B = M.Batch(_params(M), M.PREC_MIXED, lanes=4)
    hs, ref = [], []
    for cells, rho, temp, seed in SYSTEMS:
        xo, vo, fo, _ = _oracle((cells, rho, temp, seed), 30)
        L = np.array(cells) * (4.0 / rho) ** (1.0 / 3.0)
        n = len(xo)
        hs.append(B.add(np.mod(xo, L), np.zeros(n, np.int32), [0, 0, 0], L,
                        v=vo, id=np.arange(1, n + 1, dtype=np.int32)))
        ref.append(fo)
    B.run(0)
    for h, fo in zip(hs, ref):
        _, _, f, _, _ = B.download(h)
        err = np.linalg.norm(f - fo) / np.linalg.norm(fo)
        assert np.abs(fo).max() > 1.0  # a liquid, not the lattice
        assert err < 1e-5, err


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_alone_equals_in_mix(gpu, prec):
    M = gpu
    P = M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED
    A = SYSTEMS[1]
    alone = M.Batch(_params(M), P, lanes=2)
    ha = _add(M, alone, A)
    for k in (20, 20, 7):
        alone.run(k)
    mix = M.Batch(_params(M), P, lanes=2)
    hb = _add(M, mix, SYSTEMS[0])
    hm = _add(M, mix, A)
    _add(M, mix, SYSTEMS[2])
    mix.run(20)
    # one member leaves before A (A moves in the flat range), one joins
    mix.remove(hb)
    hd = _add(M, mix, SYSTEMS[3])
    assert mix.info(hd)["state"] == "waiting"
    mix.run(20)
    assert mix.info(hd)["state"] == "running" and mix.info(hd)["step"] == 20
    assert mix.info(hb)["state"] == "unknown"
    mix.run(7)
    assert mix.info(hm)["step"] == 47
    ra, rm = alone.download(ha), mix.download(hm)
    for a, b in zip(ra, rm):
        assert np.array_equal(a, b)
    assert np.array_equal(alone.thermo(ha), mix.thermo(hm))
    # the joined system's first row is its step 0
    assert mix.thermo(hd)[0, 0] == 0


def test_batch_join_waits_for_rebuild_boundary(gpu):
    M = gpu
    B = M.Batch(_params(M), M.PREC_FP64, lanes=1)
    _add(M, B, SYSTEMS[0])
    B.run(7)
    h = _add(M, B, SYSTEMS[1])
    B.run(13)  # starts at step 7: not a boundary
    assert B.info(h)["state"] == "waiting"
    B.run(1)  # starts at step 20
    assert B.info(h)["state"] == "running" and B.info(h)["step"] == 1


def test_batch_autotuned_lanes_run(gpu):
    M = gpu
    B = M.Batch(_params(M), M.PREC_MIXED, lanes=0)
    hs = [_add(M, B, s) for s in SYSTEMS]
    B.run(40)
    c = B.counts()
    assert c["lanes"] in (1, 2, 4, 8)
    for h in hs:
        th = B.thermo(h)
        e = th[:, 3]
        assert np.abs(e - e[0]).max() < 1e-3 * abs(e[0])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_graph_replays_equal_eager(gpu, prec):
    """The plain steps between rebuilds run as replays of a captured graph
    (two steps per replay, captured again after every rebuild).  Runs of
    uneven length, a system joining and one leaving, and host reads between
    the runs: every system's x, v, f and thermo rows are bit-identical to the
    same batch launched kernel by kernel, after every run."""
    M = gpu
    P = M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED
    g = M.Batch(_params(M, thermo=10), P, lanes=2, graphs=True)
    e = M.Batch(_params(M, thermo=10), P, lanes=2, graphs=False)
    hg = [_add(M, g, s) for s in SYSTEMS[:3]]
    he = [_add(M, e, s) for s in SYSTEMS[:3]]
    for k, change in ((27, None), (13, "join"), (40, "leave"), (23, None)):
        if change == "join":
            hg.append(_add(M, g, SYSTEMS[3]))
            he.append(_add(M, e, SYSTEMS[3]))
        elif change == "leave":
            g.remove(hg.pop(0))
            e.remove(he.pop(0))
        g.run(k)
        e.run(k)
        for a, b in zip(hg, he):
            # a join takes effect at the next run that starts on a rebuild
            assert g.info(a)["state"] == e.info(b)["state"]
            if g.info(a)["state"] != "running":
                continue
            for u, w in zip(g.download(a), e.download(b)):
                assert np.array_equal(u, w)
            assert np.array_equal(g.thermo(a), e.thermo(b))
    assert all(g.info(a)["state"] == "running" for a in hg)
    assert g.graph_steps() > 40
    assert e.graph_steps() == 0


def test_batch_graphs_two_types(gpu):
    """Two atom types (masses 1 and 2, their own cross coefficients): the
    coefficient tables go to the device once, so the plain steps run as
    graph replays here too, and every system's x, v, f and thermo rows are
    bit-identical to the batch launched kernel by kernel."""
    M = gpu

    def params():
        p = _params(M, thermo=10)
        p.ntypes = 2
        p.mass[1] = 2.0
        p.ncoeff = 3
        for k, (a, b, eps, sig) in enumerate(((1, 1, 1.0, 1.0), (1, 2, 0.8, 1.1),
                                              (2, 2, 1.2, 0.95))):
            p.coeff_i[k], p.coeff_j[k] = a, b
            p.coeff_eps[k], p.coeff_sigma[k], p.coeff_cut[k] = eps, sig, 2.5
        return p

    g = M.Batch(params(), M.PREC_FP64, lanes=2, graphs=True)
    e = M.Batch(params(), M.PREC_FP64, lanes=2, graphs=False)
    hs = []
    for cells, rho, temp, seed in SYSTEMS[:2]:
        x, v, t, i, lo, hi = M.fcc_system(cells, rho, temp, seed)
        t = t.copy()
        t[::3] = 1
        hs.append((g.add(x, t, lo, hi, v=v, id=i), e.add(x, t, lo, hi, v=v, id=i)))
    for k in (27, 30):
        g.run(k)
        e.run(k)
        for a, b in hs:
            for u, w in zip(g.download(a), e.download(b)):
                assert np.array_equal(u, w)
            assert np.array_equal(g.thermo(a), e.thermo(b))
    assert g.graph_steps() > 20
    assert e.graph_steps() == 0


def test_batch_precision_switch_with_graphs(gpu):
    """Switching precision between runs takes effect at once, also for the
    steps that replay captured graphs: a batch with graphs and one launched
    kernel by kernel, both switched FP64 -> mixed -> FP64 mid-way between
    rebuilds, stay bit-identical, and differ from a batch that stays FP64."""
    M = gpu
    g = M.Batch(_params(M, thermo=10), M.PREC_FP64, lanes=2, graphs=True)
    e = M.Batch(_params(M, thermo=10), M.PREC_FP64, lanes=2, graphs=False)
    s = M.Batch(_params(M, thermo=10), M.PREC_FP64, lanes=2, graphs=True)
    hg = [_add(M, g, sp) for sp in SYSTEMS[:2]]
    he = [_add(M, e, sp) for sp in SYSTEMS[:2]]
    hs = [_add(M, s, sp) for sp in SYSTEMS[:2]]
    for k, prec in ((25, None), (12, M.PREC_MIXED), (14, M.PREC_FP64), (9, None)):
        if prec is not None:
            g.set_precision(prec)
            e.set_precision(prec)
        g.run(k)
        e.run(k)
        s.run(k)
        for a, b in zip(hg, he):
            for u, w in zip(g.download(a), e.download(b)):
                assert np.array_equal(u, w)
            assert np.array_equal(g.thermo(a), e.thermo(b))
    assert g.graph_steps() > 20
    xg = g.download(hg[0])[0]
    xs = s.download(hs[0])[0]
    assert not np.array_equal(xg, xs)
