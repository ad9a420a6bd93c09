"""read_data / write_data around the timestep loop (cabanamd_impl.h:186-188,
430-431) on the GPU: a run started from a data file holding the lattice
deck's initial state reproduces the lattice run bit for bit (FP64, one
lane per atom), and write_data leaves the final state in the file."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _by_id(x, v, t, i, n):
    o = np.argsort(i[:n], kind="stable")
    return x[:n][o], v[:n][o], t[:n][o], i[:n][o]


def test_read_data_run_matches_lattice_run(gpu, tmp_path):
    M = gpu
    p = M.lattice_params(ncells=(6, 6, 6), thermo=10, nsteps=30)
    a = M.CbnMD(p, precision=M.PREC_FP64, lanes=1)
    a.init()
    n = a.n_local
    x0, v0, t0, i0 = _by_id(*[a.arrays()[k] for k in (0, 1, 3, 4)], n)
    L = p.lattice_constant * 6
    data = os.path.join(tmp_path, "start.data")
    M.write_data(data, x0, v0, t0, i0, [0, 0, 0], [L, L, L], 1)
    a.run(30)
    q = M.lattice_params(ncells=(6, 6, 6), thermo=10, nsteps=30)
    q.data_in = data.encode()
    out = os.path.join(tmp_path, "end.data")
    q.data_out = out.encode()
    b = M.CbnMD(q, precision=M.PREC_FP64, lanes=1)
    b.init()
    b.run(30)
    xa, va, ta, ia = _by_id(*[a.arrays()[k] for k in (0, 1, 3, 4)], n)
    xb, vb, tb, ib = _by_id(*[b.arrays()[k] for k in (0, 1, 3, 4)], n)
    assert np.array_equal(ia, ib)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    assert np.array_equal(a.thermo(), b.thermo())
    d = M.read_data(out)
    assert np.array_equal(d["id"], ib) and np.array_equal(d["v"], vb)
    # positions come back wrapped into the box
    assert np.all((d["x"] >= 0) & (d["x"] < L))
    dx = d["x"] - xb
    assert np.abs(dx - L * np.round(dx / L)).max() < 1e-12
