"""NNP force on the GPU against the FP64 NNP oracle, through the C ABI.

Symmetry functions and forces run in FP32 and the network layers in
split-BF16 MFMA (CMD_PREC_MIXED: each FP32 operand as three BF16 parts, six
part products per layer product, FP32 accumulation) or FP32 on the VALU
(CMD_PREC_FP64), so the bars are relative:
  G:        max |dG| <= 2e-5 * max |G| per function
  E_i:      1e-4 of the energy scale
  forces:   ||dF||/||F|| <= 1e-3, and <= 1e-5 on the benchmark model (the
            bar the batched and served forms are held to)
A single BF16 product would carry its 8-bit mantissa (~4e-3 relative per
operand); the split brings the layers to FP32 accuracy.
"""
import numpy as np
import pytest

import pyoracle as O
from nnp_models import FIXTURE, random_model, to_oracle

pytestmark = pytest.mark.gpu


def _grid_model(seed=3, rc=5.0, ang_rs=0.0):
    """random_model without its element-specific angular function, so the
    functions form the grid the kernels evaluate; ang_rs moves every angular
    function's rs (0: the kernels' shared-argument path, else the general
    one)."""
    m = random_model(seed=seed, rc=rc)
    for e in range(2):
        keep = [k for k, s in enumerate(m["sf"][e]) if not (s[0] == 3 and s[1] >= 0)]
        m["sf"][e] = [m["sf"][e][k] if m["sf"][e][k][0] == 2 else
                      m["sf"][e][k][:4] + (ang_rs,) + m["sf"][e][k][5:]
                      for k in keep]
        m["gmin"][e] = m["gmin"][e][keep]
        m["gmax"][e] = m["gmax"][e][keep]
        m["W"][e][0] = m["W"][e][0][:, keep]
    return m


def _engine_model(M, m):
    em = M.NNPModel.create(m["nelem"], m["hidden"], m["act"])
    em.set_scaling(m["scale"], m["smin"], m["smax"])
    for e in range(m["nelem"]):
        for k, s in enumerate(m["sf"][e]):
            em.add_sf(e, *s, gmin=float(m["gmin"][e][k]),
                      gmax=float(m["gmax"][e][k]))
        for l in range(len(m["W"][e])):
            em.set_layer(e, l, m["W"][e][l], m["b"][e][l])
        em.set_atom_energy(e, m["atom_energy"][e])
    return em


def _system(M, x, t, L, cut, ntypes=2):
    s = M.System(ntypes)
    s.create_domain([0, 0, 0], [L[0], L[1], L[2]], cut)
    s.upload(x, None, None, t, np.arange(1, len(x) + 1, dtype=np.int32))
    comm = M.Comm(s, cut)
    comm.exchange_halo()
    nl = M.NeighborVerlet(cut, False, 100, lanes=1)
    nl.create(s)
    return s, comm, nl


def _run(M, s, comm, nl, F, m, ld=96):
    F.compute(s, nl)
    comm.update_force()
    xg, _, fg, tg, _ = s.download()
    G, D, E = F.atomic(s, ld)
    off, idx = nl.get()
    n = s.N_local
    Mo = to_oracle(O, m)
    eo, Eio, Go, Do, fo = Mo.compute(xg, tg, n, off, idx, ld=ld)
    send, counts = comm.ghost_lists()
    O.ghost_fold(fo, n, send, counts)
    return dict(G=G, D=D, E=E, f=fg[:n], Go=Go, Do=Do, Eo=Eio, fo=fo[:n], n=n,
                eo=eo)


def _check(r, nsf, prec):
    G, Go = r["G"][:, :nsf], r["Go"][:, :nsf]
    scale = np.maximum(np.abs(Go).max(axis=0), 1e-12)
    assert (np.abs(G - Go) / scale).max() < 2e-5
    escale = max(np.abs(r["Eo"]).max(), 1e-3)
    etol = 1e-4
    assert np.abs(r["E"] - r["Eo"]).max() / escale < etol
    ftol = 1e-3
    err = np.linalg.norm(r["f"] - r["fo"]) / np.linalg.norm(r["fo"])
    assert err < ftol, err
    assert np.abs(r["fo"]).max() > 1e-6  # forces are not trivially zero
    return err


@pytest.mark.parametrize("ang_rs", [0.0, 0.4])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_nnp_random_model_periodic(gpu, prec, ang_rs):
    M = gpu
    m = _grid_model(ang_rs=ang_rs)
    x, t, _ = M.nnp_lattice((6, 6, 6), spacing=2.0, jitter=0.1, seed=11)
    L = [12.0] * 3
    s, comm, nl = _system(M, x, t, L, 5.3)
    F = M.ForceNNP(_engine_model(M, m),
                   M.PREC_FP64 if prec == "fp32" else M.PREC_MIXED)
    r = _run(M, s, comm, nl, F, m)
    err = _check(r, len(m["sf"][0]), prec)
    print(f"random model {prec}, angular rs {ang_rs}: force rel L2 {err:.2e}")
    # net force of a periodic system vanishes
    assert np.abs(r["f"].sum(axis=0)).max() < 1e-3 * np.abs(r["f"]).max()


def test_nnp_work_counts(gpu):
    """The work model's term counts (cmd_force_nnp_work): neighbours of each
    local atom inside the model's cutoff and neighbour pairs (j, k) with r_jk
    inside it, summed over the atoms, equal a brute-force FP64 count on the
    same positions (pairs within 1e-4 of the cutoff, which the kernels'
    FP32 distances may class either way, are the only slack)."""
    M = gpu
    rc = 5.0
    m = _grid_model(rc=rc)
    x, t, _ = M.nnp_lattice((6, 6, 6), spacing=2.0, jitter=0.1, seed=11)
    L = [12.0] * 3
    s, comm, nl = _system(M, x, t, L, 5.3)
    F = M.ForceNNP(_engine_model(M, m), M.PREC_FP64)
    F.compute(s, nl)
    w = F.work()
    xg = s.download()[0]
    n = s.N_local
    assert w["atoms"] == n
    nb = tri = slack_nb = slack_tri = 0
    for i in range(n):
        d = np.linalg.norm(xg - xg[i], axis=1)
        d[i] = np.inf
        js = np.nonzero(d < rc)[0]
        nb += len(js)
        slack_nb += int(np.sum(np.abs(d - rc) < 1e-4))
        p = xg[js]
        djk = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
        iu = np.triu_indices(len(js), 1)
        tri += int(np.sum(djk[iu] < rc))
        slack_tri += int(np.sum(np.abs(djk[iu] - rc) < 1e-4))
    assert nb > 0 and tri > 0
    assert abs(w["neighbours"] - nb) <= slack_nb, (w["neighbours"], nb, slack_nb)
    assert abs(w["triplets"] - tri) <= slack_tri + 60 * slack_nb, (w["triplets"], tri)
    # every term count enters the flops with a positive weight
    assert w["flop_symmetry_functions"] > 20.0 * nb
    assert w["flop_force_chain"] > w["flop_symmetry_functions"]


def test_nnp_benchmark_model_after_scaling_fit(gpu):
    M = gpu
    x, t, _ = M.nnp_lattice((8, 8, 8), spacing=2.5, jitter=0.1, seed=5)
    L = [20.0] * 3
    s, comm, nl = _system(M, x, t, L, 6.3)
    em = M.NNPModel.synthetic(seed=7, rc=6.0)
    F = M.ForceNNP(em, M.PREC_FP64)
    F.fit_scaling(s, nl)
    m = F.model().describe(2)
    m.update(act="ttl", smin=0.0, smax=1.0, scale=True)
    # the fitted ranges cover every atom's functions
    assert all(np.all(m["gmax"][e] >= m["gmin"][e]) for e in range(2))
    r = _run(M, s, comm, nl, F, m)
    err32 = _check(r, 70, "fp32")
    F.set_precision(M.PREC_MIXED)
    r = _run(M, s, comm, nl, F, m)
    err = _check(r, 70, "bf16")
    print(f"benchmark model: force rel L2 {err32:.2e} (FP32 layers), "
          f"{err:.2e} (split-BF16 MFMA layers)")
    assert err32 < 1e-5 and err < 1e-5


def test_nnp_reference_fixture(gpu):
    """input/nnp on the in.nnp crystal (fcc 3.52 A, 3x3x3 cells).  The
    model's symmetry functions are in Bohr (rc 7.37 = 3.9 A), so positions
    are converted with in.nnp's cflength 1.8897261328 (the reference passes
    raw positions, which saturates this network); list at 7.37 + 1.0.  A
    perfect crystal gives equal atomic energies and zero forces; a jittered
    copy is checked against the oracle."""
    M = gpu
    a = 3.52 * 1.8897261328
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    x = np.array([(np.array([i, j, k]) + b) * a for k in range(3)
                  for j in range(3) for i in range(3) for b in base])
    t = np.zeros(len(x), np.int32)
    em = M.NNPModel.load_n2p2(FIXTURE, ["Ni"], [28])
    m = em.describe(1)
    m.update(act="ttl", smin=0.0, smax=1.0, scale=True)
    L = [3 * a] * 3
    cut = 7.3699319 + 1.0
    s, comm, nl = _system(M, x, t, L, cut, ntypes=1)
    F = M.ForceNNP(em, M.PREC_FP64)
    r = _run(M, s, comm, nl, F, m)
    assert np.allclose(r["E"], r["E"][0], rtol=1e-5, atol=1e-6)
    assert np.abs(r["f"]).max() < 1e-4
    # scaled inputs of the crystal lie inside the fixture's training range
    gs = (r["Go"][0, :27] - m["gmin"][0]) / (m["gmax"][0] - m["gmin"][0])
    assert np.all(gs > -0.5) and np.all(gs < 1.5)
    rng = np.random.default_rng(1)
    xj = np.mod(x + rng.uniform(-0.2, 0.2, x.shape), L[0])
    s, comm, nl = _system(M, xj, t, L, cut, ntypes=1)
    r = _run(M, s, comm, nl, F, m)
    _check(r, 27, "fp32")
    # ForceNNP::compute_energy's assembly (impl.h:101-130)
    e = F.compute_energy(s, nl)
    ref = (r["Eo"].sum() + len(x) * -0.97) * 27.211384021355236
    assert abs(e - ref) < 1e-4 * abs(ref)


def _fitted_benchmark_model(M, x, t, L, seed=7):
    s, comm, nl = _system(M, x, t, L, 6.3)
    F = M.ForceNNP(M.NNPModel.synthetic(seed=seed, rc=6.0), M.PREC_FP64)
    F.fit_scaling(s, nl)
    return F.model()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_nnp_md_loop_conserves_energy(gpu, prec):
    """CbnMD with the NNP force (cabanamd_impl.h:285-399, unfused order):
    3000 atoms, 100 steps of dt 0.0005 with a rebuild every 20; the total
    energy (sum of atomic energies + kinetic) stays within 1e-3 relative
    (FP32 symmetry functions, forces)."""
    M = gpu
    x, t, ids = M.nnp_lattice((10, 15, 20), spacing=2.5, jitter=0.1, seed=5)
    L = [25.0, 37.5, 50.0]
    model = _fitted_benchmark_model(M, x, t, L)
    rng = np.random.default_rng(3)
    v = rng.normal(0, np.sqrt(0.05), x.shape)
    v -= v.mean(axis=0)
    p = M.nnp_params(nsteps=100, thermo=10)
    md = M.CbnMD(p, precision=M.PREC_FP64 if prec == "fp32" else M.PREC_MIXED)
    md.set_nnp_model(model)
    md.set_initial_state(x, t, [0, 0, 0], L, v=v, id=ids)
    md.init()
    md.run(100)
    th = md.thermo()
    n = len(x)
    pe = th[:, 2] * n / 27.211384021355236  # the reference's Hartree factor
    ke = (th[:, 3] - th[:, 2]) * n
    e = pe + ke
    drift = np.abs(e - e[0]).max() / np.abs(e).max()
    print(f"NNP MD {prec}: energy drift {drift:.2e} over 100 steps")
    assert th.shape[0] == 11
    assert drift < 1e-3
