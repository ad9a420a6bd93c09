"""NNP model handling on the host (no GPU): the engine's n2p2 reader against
an independent parser of the reference fixture, the seeded benchmark model,
the jittered lattice, and the checks that reject unsupported models."""
import numpy as np
import pytest

import cabanamd_amd as M
from nnp_models import FIXTURE, read_n2p2


def test_engine_reads_the_fixture_like_the_independent_parser():
    ref = read_n2p2()
    got = M.NNPModel.load_n2p2(FIXTURE, ["Ni"], [28]).describe(1)
    assert got["hidden"] == ref["hidden"] == [24, 24]
    assert got["sf"][0] == ref["sf"][0]
    assert np.array_equal(got["gmin"][0], ref["gmin"][0])
    assert np.array_equal(got["gmax"][0], ref["gmax"][0])
    for l in range(3):
        assert np.array_equal(got["W"][0][l], ref["W"][0][l])
        assert np.array_equal(got["b"][0][l], ref["b"][0][l])
    assert got["atom_energy"] == ref["atom_energy"] == [-0.97]
    nparams = sum(w.size + b.size for w, b in zip(got["W"][0], got["b"][0]))
    assert nparams == 1297


def test_missing_model_files_are_an_error():
    with pytest.raises(M.CmdError, match="input.nn"):
        M.NNPModel.load_n2p2("/nonexistent", ["Ni"], [28])


def test_synthetic_benchmark_model():
    m = M.NNPModel.synthetic(seed=5, rc=6.0).describe(2)
    for e in range(2):
        sf = m["sf"][e]
        assert len(sf) == 70
        rad = [s for s in sf if s[0] == 2]
        ang = [s for s in sf if s[0] == 3]
        assert len(rad) == 30 and len(ang) == 40
        assert sorted({s[1] for s in rad}) == [0, 1]
        etas = sorted({s[3] for s in rad})
        assert len(etas) == 15 and np.isclose(etas[0], 0.01) and np.isclose(etas[-1], 2.0)
        assert sorted({s[6] for s in ang}) == [1.0, 2.0, 4.0, 16.0]
        assert sorted({s[5] for s in ang}) == [-1.0, 1.0]
        assert all(s[7] == 6.0 for s in sf)
        assert [w.shape for w in m["W"][e]] == [(25, 70), (25, 25), (1, 25)]
        # N(0, 1) / sqrt(fan_in)
        assert abs(m["W"][e][0].std() * np.sqrt(70) - 1.0) < 0.1
    # seeded: same seed, same weights; another seed differs
    again = M.NNPModel.synthetic(seed=5, rc=6.0).describe(2)
    assert np.array_equal(again["W"][1][1], m["W"][1][1])
    other = M.NNPModel.synthetic(seed=6, rc=6.0).describe(2)
    assert not np.array_equal(other["W"][1][1], m["W"][1][1])


def test_jittered_lattice():
    x, t, i = M.nnp_lattice((10, 15, 20), spacing=2.5, jitter=0.1, seed=3)
    assert len(x) == 3000 and sorted(i) == list(range(1, 3001))
    assert (t == 1).sum() == 1000 and (t == 0).sum() == 2000
    # every coordinate within +-0.1 s of its site
    s = 2.5
    frac = x / s - 0.5
    assert np.abs(frac - np.round(frac)).max() <= 0.1 + 1e-12
    # nearest-neighbour separation >= 0.8 s on a sample
    d = np.linalg.norm(x[:300, None, :] - x[None, :, :], axis=-1)
    d[np.arange(300), np.arange(300)] = 1e9
    assert d.min() >= 0.8 * s - 1e-12


def test_non_grid_symmetry_functions_are_rejected():
    m = M.NNPModel.create(2, [4, 4], "ttl")
    for e in range(2):
        m.add_sf(e, 2, 0, -1, 0.1, 0.0, 0.0, 0.0, 5.0)
        m.add_sf(e, 2, 1, -1, 0.1, 0.0, 0.0, 0.0, 5.0)
        m.add_sf(e, 2, 0, -1, 0.3, 0.0, 0.0, 0.0, 5.0)  # (0.3, e=1) missing
        for l, (o, i) in enumerate([(4, 3), (4, 4), (1, 4)]):
            m.set_layer(e, l, np.zeros((o, i)), np.zeros(o))
    with pytest.raises(M.CmdError, match="grid"):
        M.ForceNNP(m)
