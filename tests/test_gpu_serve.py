"""MD request server (cmd_server) on the GPU, through the C ABI.

Continuous batching must not change results: a job's result alone and in a
mix of jobs that arrive, wait for slots, join and leave around it is the
same -- bit-identical for LJ (deterministic kernels and reductions), within
1e-9 for NNP (its force kernel accumulates neighbour reactions with FP64
atomics, whose order varies at the last bit).  An LJ job also matches the
FP64 oracle's MD loop from the same seeded state, and admission respects
the device-memory budget.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _lj_params(M):
    return M.lattice_params(ncells=(5, 5, 5), thermo=100, every=20)


@pytest.fixture(scope="module")
def nnp_model(gpu):
    M = gpu
    x, t, _ = M.nnp_lattice((8, 8, 8), spacing=2.5, jitter=0.1, seed=5)
    L = [20.0] * 3
    s = M.System(2)
    s.create_domain([0, 0, 0], L, 6.3)
    s.upload(np.mod(x, 20.0), None, None, t,
             np.arange(1, len(x) + 1, dtype=np.int32))
    comm = M.Comm(s, 6.3)
    comm.exchange_halo()
    nl = M.NeighborVerlet(6.3, False, 100, lanes=1)
    nl.create(s)
    F = M.ForceNNP(M.NNPModel.synthetic(seed=7, rc=6.0), M.PREC_FP64)
    F.fit_scaling(s, nl)
    return F.model()


def _server(M, model, **kw):
    return M.Server(_lj_params(M), M.nnp_params(thermo=100), model,
                    precision=M.PREC_MIXED, lanes=2, **kw)


LJ_A = ((5, 6, 5), 0.8, 1.4, 31)
NNP_B = ((10, 10, 11), 0.05, 32)


def _submit_lj(M, S, spec, nsteps, arrival=0.0):
    x, v, t, i, lo, hi = M.fcc_system(*spec)
    return S.submit(S.LJ, x, t, lo, hi, nsteps, v=v, id=i, arrival=arrival)


def _submit_nnp(M, S, spec, nsteps, arrival=0.0):
    cells, temp, seed = spec
    x, v, t, i, lo, hi = M.nnp_system(cells, temp=temp, seed=seed)
    return S.submit(S.NNP, x, t, lo, hi, nsteps, v=v, id=i, arrival=arrival)


def test_server_results_same_alone_and_in_mix(gpu, nnp_model):
    M = gpu
    alone = _server(M, nnp_model)
    a = _submit_lj(M, alone, LJ_A, 57)
    b = _submit_nnp(M, alone, NNP_B, 43)
    alone.drain(300)
    ra, rb = alone.result(a), alone.result(b)
    # a mix with three slots: the two jobs wait for others to finish, join
    # mid-run of the rest and leave before them
    mix = _server(M, nnp_model, max_jobs=3)
    _submit_lj(M, mix, ((6, 6, 6), 0.75, 1.0, 41), 30)
    _submit_nnp(M, mix, ((10, 10, 10), 0.04, 42), 25)
    _submit_lj(M, mix, ((5, 5, 5), 0.85, 1.8, 43), 90)
    ma = _submit_lj(M, mix, LJ_A, 57)
    mb = _submit_nnp(M, mix, NNP_B, 43)
    _submit_lj(M, mix, ((7, 6, 6), 0.9, 0.9, 44), 44)
    mix.drain(300)
    for j in range(6):
        assert mix.job(j)["state"] == mix.DONE
    # the late jobs had to wait for a slot
    assert mix.job(ma)["admit"] >= min(mix.job(k)["done"] for k in range(3))
    xa, va, ia, tha = mix.result(ma)
    assert np.array_equal(xa, ra[0]) and np.array_equal(va, ra[1])
    assert np.array_equal(ia, ra[2])
    assert np.array_equal(tha[-1], ra[3][-1]) and tha[-1, 0] == 57
    xb, vb, ib, thb = mix.result(mb)
    assert np.array_equal(ib, rb[2])
    assert np.abs(xb - rb[0]).max() < 1e-9
    assert np.abs(vb - rb[1]).max() < 1e-9
    assert thb[-1, 0] == 43
    assert np.allclose(thb[-1, 1:], rb[3][-1, 1:], rtol=1e-9)
    st = mix.stats()
    assert st["done"] == 6 and st["running"] == 0 and st["bytes"] == 0


def test_server_lj_job_matches_oracle(gpu, nnp_model):
    M = gpu
    S = M.Server(_lj_params(M), precision=M.PREC_FP64, lanes=1)
    j = _submit_lj(M, S, LJ_A, 45)
    S.drain(120)
    x, v, _, th = S.result(j)
    cells, rho, temp, seed = LJ_A
    md = O.MD(O.lj_params(ncells=cells, rho=rho, temp=temp, seed=seed,
                          gaussian=True, thermo=5, every=20))
    md.init()
    md.run(45)
    xo, vo, _, _, io = md.arrays()
    n = md.n_local
    o = np.argsort(io[:n], kind="stable")
    L = np.array(cells) * (4.0 / rho) ** (1.0 / 3.0)
    d = x - xo[:n][o]
    assert np.abs(d - L * np.round(d / L)).max() < 1e-9
    assert np.abs(v - vo[:n][o]).max() < 1e-9
    tho = md.thermo()
    assert np.allclose(th[-1, 1:], tho[tho[:, 0] == 45][0, 1:], rtol=1e-10)


def test_server_admission_respects_budget(gpu, nnp_model):
    M = gpu
    x, v, t, i, lo, hi = M.fcc_system((6, 6, 6), 0.8, 1.0, 51)
    # room for one job of 864 atoms at a time
    S = M.Server(_lj_params(M), hbm_budget=len(x) * 800 + 1000,
                 precision=M.PREC_FP64, lanes=1)
    j0 = S.submit(S.LJ, x, t, lo, hi, 25, v=v, id=i)
    j1 = S.submit(S.LJ, x, t, lo, hi, 25, v=v, id=i)
    S.drain(120)
    a, b = S.job(j0), S.job(j1)
    assert b["admit"] >= a["done"]
    # the same job twice gives the same result
    r0, r1 = S.result(j0), S.result(j1)
    assert np.array_equal(r0[0], r1[0]) and np.array_equal(r0[1], r1[1])
    with pytest.raises(M.CmdError):
        S.submit(S.LJ, np.zeros((2000, 3)), np.zeros(2000, np.int32), lo, hi, 10)


def test_server_nnp_job_conserves_energy(gpu, nnp_model):
    M = gpu
    S = _server(M, nnp_model)
    j = _submit_nnp(M, S, NNP_B, 200)
    S.drain(300)
    _, _, _, th = S.result(j)
    n = S.job(j)["n"]
    pe = th[:, 2] * n / 27.211384021355236
    ke = (th[:, 3] - th[:, 2]) * n
    e = pe + ke
    assert th[0, 0] == 0 and th[-1, 0] == 200
    assert np.abs(e - e[0]).max() / np.abs(e).max() < 1e-2
