"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the pinned oracle).

CPU: the oracle still reproduces every fixture bit for bit, and the fixture
lists satisfy the tstNeighbor.hpp checks.  GPU: the HIP engine reproduces the
256-atom trajectories from the fixtures alone (no oracle on the device box).
"""
import os

import numpy as np
import pytest

import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.mark.parametrize("half", [False, True])
def test_oracle_reproduces_lj256(half):
    g = _load("lj256_half.npz" if half else "lj256_full.npz")
    md = O.MD(O.lj_params((4, 4, 4), half=half, thermo=5, nsteps=20))
    md.init()
    md.run(20)
    x, v, f, t, ids = md.arrays()
    n = md.n_local
    o = np.argsort(ids[:n])
    assert np.array_equal(x[:n][o], g["x20"])
    assert np.array_equal(v[:n][o], g["v20"])
    assert np.array_equal(f[:n][o], g["f20"])
    assert np.array_equal(md.thermo(), g["thermo"])


def test_oracle_reproduces_rng():
    g = _load("rng.npz")
    for k in range(len(g["seed"])):
        s = O.rng_reset(int(g["ibase"][k]), g["coord"][k])
        assert s == g["seed"][k]
        for m in range(3):
            u, s = O.rng_uniform(s)
            assert u == g["uniform"][k, m]


def test_neighbor_fixture_against_cell_list_build():
    g = _load("neigh_tst.npz")
    x, n_local, rc = g["x"], int(g["n_local"]), float(g["rc"])
    off, idx = O.neigh_build(x, n_local, rc, half=False)
    assert np.array_equal(off, g["offsets"])
    for i in range(n_local):
        assert sorted(idx[off[i]:off[i + 1]]) == \
            sorted(g["idx"][g["offsets"][i]:g["offsets"][i + 1]])


@pytest.mark.gpu
@pytest.mark.parametrize("half", [False, True])
def test_engine_reproduces_lj256(gpu, half):
    """FP64, one lane per atom: the full-list trajectory is bit-identical to
    the fixture; the half list (atomic scatter order) within 1e-10."""
    M = gpu
    g = _load("lj256_half.npz" if half else "lj256_full.npz")
    md = M.CbnMD(M.lattice_params((4, 4, 4), half=half, thermo=5, nsteps=20),
                 precision=M.PREC_FP64, lanes=1)
    md.init()
    x, v, f, t, ids = md.arrays()
    n = md.n_local
    o = np.argsort(ids[:n])
    assert np.array_equal(x[:n][o], g["x0"])
    assert np.array_equal(v[:n][o], g["v0"])
    md.run(20)
    x, v, f, t, ids = md.arrays()
    o = np.argsort(ids[:n])
    th = md.thermo()[:, :4]
    if half:
        assert np.abs(x[:n][o] - g["x20"]).max() < 1e-10
        assert np.abs(v[:n][o] - g["v20"]).max() < 1e-10
        assert np.allclose(th, g["thermo"], rtol=1e-10, atol=1e-12)
    else:
        assert np.array_equal(x[:n][o], g["x20"])
        assert np.array_equal(v[:n][o], g["v20"])
        assert np.array_equal(f[:n][o], g["f20"])
        assert np.allclose(th, g["thermo"], rtol=1e-12, atol=1e-14)
