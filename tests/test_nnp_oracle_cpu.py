"""Pins the FP64 NNP oracle (oracle/nnp.cpp).

The reference's NNP arithmetic lives in n2p2 v2.0.1, which is absent, so the
oracle is checked against what the mathematics fixes: an independent
brute-force evaluation of every symmetry function, forces as the exact
negative gradient of the energy (central differences), invariance under
translation, rotation and atom relabelling, zero net force, and the
reference fixture (input/nnp) read by an independent parser.
"""
import numpy as np
import pytest

import pyoracle as O
from nnp_models import cluster, random_model, read_n2p2, sf_brute, to_oracle


def _brute_list(x, rc):
    off, idx = O.neigh_brute(x, len(x), rc, half=False)
    return off, idx


def test_symmetry_functions_match_brute_force():
    m = random_model()
    M = to_oracle(O, m)
    x, t = cluster(14, box=4.5)
    off, idx = _brute_list(x, 4.5)
    G = M.symmetry_functions(x, t, len(x), off, idx)
    ref = sf_brute(m, x, t)
    for i in range(len(x)):
        n = len(ref[i])
        assert np.allclose(G[i, :n], ref[i], rtol=1e-12, atol=1e-14)
    # angular functions are present and non-trivial
    assert np.abs(G[:, 6:]).max() > 1e-3


def test_forces_are_minus_energy_gradient():
    m = random_model()
    M = to_oracle(O, m)
    x, t = cluster(20)
    off, idx = _brute_list(x, 4.6)  # list margin covers the displacements
    e, Ei, G, dG, f = M.compute(x, t, len(x), off, idx)
    assert np.isclose(e, Ei.sum())
    h = 1e-5
    fd = np.zeros_like(f)
    for i in range(len(x)):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, c] += h
            xm[i, c] -= h
            fd[i, c] = -(M.compute(xp, t, len(x), off, idx)[0] -
                         M.compute(xm, t, len(x), off, idx)[0]) / (2 * h)
    assert np.abs(fd - f).max() < 1e-7 * max(1.0, np.abs(f).max())
    assert np.abs(f).max() > 1e-3


def test_invariances_and_zero_net_force():
    m = random_model(seed=11)
    M = to_oracle(O, m)
    x, t = cluster(22, seed=9)
    off, idx = _brute_list(x, 4.6)
    e, _, _, _, f = M.compute(x, t, len(x), off, idx)
    assert np.abs(f.sum(axis=0)).max() < 1e-11 * max(1.0, np.abs(f).max())
    # translation
    e2, _, _, _, f2 = M.compute(x + np.array([1.3, -0.7, 2.1]), t, len(x),
                                off, idx)
    assert abs(e2 - e) < 1e-11 * abs(e)
    assert np.allclose(f2, f, atol=1e-11)
    # rotation
    rng = np.random.default_rng(2)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    xr = x @ q.T
    e3, _, _, _, f3 = M.compute(xr, t, len(x), off, idx)
    assert abs(e3 - e) < 1e-11 * abs(e)
    assert np.allclose(f3, f @ q.T, atol=1e-10)
    # relabelling
    p = rng.permutation(len(x))
    xp, tp = x[p], t[p]
    offp, idxp = _brute_list(xp, 4.6)
    e4, _, _, _, f4 = M.compute(xp, tp, len(x), offp, idxp)
    assert abs(e4 - e) < 1e-11 * abs(e)
    assert np.allclose(f4, f[p], atol=1e-11)


def test_zeta_one_angular_derivative():
    """zeta = 1 takes the derivative branch without the power."""
    m = random_model(seed=4)
    for e in range(2):
        m["sf"][e] = [s for s in m["sf"][e] if s[0] == 3 and s[6] == 1.0]
        m["gmin"][e] = m["gmin"][e][:len(m["sf"][e])]
        m["gmax"][e] = m["gmax"][e][:len(m["sf"][e])]
        n = len(m["sf"][e])
        m["W"][e][0] = m["W"][e][0][:, :n]
    M = to_oracle(O, m)
    x, t = cluster(10, box=3.5, seed=1)
    off, idx = _brute_list(x, 4.6)
    e0, _, _, _, f = M.compute(x, t, len(x), off, idx)
    h = 1e-5
    for i in (0, 3, 7):
        xp, xm = x.copy(), x.copy()
        xp[i, 1] += h
        xm[i, 1] -= h
        fd = -(M.compute(xp, t, len(x), off, idx)[0] -
               M.compute(xm, t, len(x), off, idx)[0]) / (2 * h)
        assert abs(fd - f[i, 1]) < 1e-7 * max(1.0, abs(f[i, 1]))


def test_reference_fixture_parses_and_runs():
    """input/nnp (Ni, 27 symmetry functions, 27-24-24-1, t t l): 1,297
    parameters, scaling rows ordered radial then angular, and the model runs
    on the in.nnp fcc crystal (lattice 3.52, 3x3x3 cells) with zero net
    force and forces equal to the energy gradient."""
    m = read_n2p2()
    assert m["nparams"] == 27 * 24 + 24 + 24 * 24 + 24 + 24 + 1 == 1297
    assert len(m["sf"][0]) == 27 and m["hidden"] == [24, 24]
    assert m["act"] == "ttl" and m["scale"]
    assert all(s[0] == 2 for s in m["sf"][0][:3])
    assert np.all(m["gmax"][0] > m["gmin"][0])
    # radial functions decrease with eta; so do their scaling ranges
    assert np.all(np.diff(m["gmax"][0][:3]) < 0)
    M = to_oracle(O, m)
    a = 3.52
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    x = np.array([(np.array([i, j, k]) + b) * a for i in range(3)
                  for j in range(3) for k in range(3) for b in base])
    t = np.zeros(len(x), np.int32)
    # periodic images through the oracle's ghost build, list at 3.9 + 1.0
    L = 3 * a
    d = O.create_domain([0, 0, 0], [L, L, L], L)
    xg, tg, ig, send, counts = O.ghost_build(x, t, np.arange(1, len(x) + 1,
                                                              dtype=np.int32),
                                             d, 4.9)
    off, idx = O.neigh_build(xg, len(x), 4.9, False)
    e, Ei, G, dG, f = M.compute(xg, tg, len(x), off, idx)
    O.ghost_fold(f, len(x), send, counts)
    # a perfect crystal: every atom sees the same environment
    assert np.allclose(Ei, Ei[0], rtol=1e-10)
    assert np.abs(f[:len(x)]).max() < 1e-9
    # scaled inputs of the crystal sit inside the training range's span
    assert np.isfinite(e)


def _numpy_network(m, e, G):
    """Independent forward pass of one element's network from raw symmetry
    functions: min-max scaling smin + (smax - smin)(G - Gmin)/(Gmax - Gmin),
    then h <- act(W h + b) per layer with W stored (out x in), tanh on the
    hidden layers and identity on the output ('ttl')."""
    h = np.asarray(G, float)
    if m["scale"]:
        h = m["smin"] + (m["smax"] - m["smin"]) * (h - m["gmin"][e]) / (
            m["gmax"][e] - m["gmin"][e])
    nl = len(m["W"][e])
    for l in range(nl):
        h = m["W"][e][l] @ h + m["b"][e][l]
        if m["act"][l] == "t":
            h = np.tanh(h)
    assert h.shape == (1,)
    return float(h[0])


def _numpy_dE_dG(m, e, G, step=1e-6):
    """dE/dG of the numpy network by central differences in each raw G."""
    out = np.zeros(len(G))
    for k in range(len(G)):
        gp, gm = np.array(G, float), np.array(G, float)
        gp[k] += step
        gm[k] -= step
        out[k] = (_numpy_network(m, e, gp) - _numpy_network(m, e, gm)) / (2 * step)
    return out


def test_network_forward_matches_numpy():
    """The oracle's atomic energies E_i equal an independent numpy forward pass
    (scaling, W h + b, tanh, tanh, linear) over the brute-force symmetry
    functions, and its dE/dG equals the numpy network's gradient.  A
    transposed W (non-square first layer), dropped scaling, a tanh on the
    output layer or a missing bias all change E_i by far more than 1e-12."""
    m = random_model(seed=21, hidden=(7, 5))
    M = to_oracle(O, m)
    x, t = cluster(16, box=4.5, seed=3)
    off, idx = _brute_list(x, 4.5)
    e, Ei, G, dG, f = M.compute(x, t, len(x), off, idx)
    ref = sf_brute(m, x, t)
    for i in range(len(x)):
        n = len(ref[i])
        want = _numpy_network(m, t[i], ref[i])
        assert abs(Ei[i] - want) <= 1e-12 * max(1.0, abs(want)), (i, Ei[i], want)
        d = _numpy_dE_dG(m, t[i], ref[i])
        assert np.allclose(dG[i, :n], d, rtol=1e-6, atol=1e-8)
    assert abs(e - Ei.sum()) <= 1e-12 * max(1.0, abs(e))
    # the checks can fail: a network with the output layer through tanh
    # differs from the oracle's linear output
    bad = dict(m, act="ttt")
    assert max(abs(_numpy_network(bad, t[i], ref[i]) - Ei[i])
               for i in range(len(x))) > 1e-6


def test_network_forward_matches_numpy_on_reference_fixture():
    """The same independent forward pass on the input/nnp model (27-24-24-1,
    scaled inputs) on a jittered fcc cluster of Ni at the in.nnp lattice
    constant 3.52: E_i from the oracle equals numpy's."""
    m = read_n2p2()
    M = to_oracle(O, m)
    rng = np.random.default_rng(8)
    a = 3.52
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    x = np.array([(np.array([i, j, k]) + b) * a for i in range(2)
                  for j in range(2) for k in range(2) for b in base])
    x = x + rng.uniform(-0.1, 0.1, x.shape)
    t = np.zeros(len(x), np.int32)
    rc = max(s[7] for s in m["sf"][0])
    off, idx = _brute_list(x, rc + 0.5)
    e, Ei, G, dG, f = M.compute(x, t, len(x), off, idx)
    ref = sf_brute(m, x, t)
    for i in range(len(x)):
        assert np.allclose(G[i, :27], ref[i], rtol=1e-10, atol=1e-12)
        want = _numpy_network(m, 0, ref[i])
        assert abs(Ei[i] - want) <= 1e-11 * max(1.0, abs(want)), (i, Ei[i], want)
