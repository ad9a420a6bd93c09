"""The LJ timestep at its benchmark size: fcc rho 0.8442, 80^3 cells =
2,048,000 atoms, rc 2.5, skin 0.3, full list rebuilt every 20 steps,
dt 0.005, two lanes per atom on tiles six cells tall (the layout the
autotuner picks for this workload); mixed precision and FP64 with the full
list, FP64 with the half list.  At this size
the checks are properties that do not depend on the size:
- the step with the ghost refresh fused into it equals the separate
  refresh pass bit for bit (positions, velocities, forces, thermo rows;
  the half list always takes the separate pass, and its reactions are
  summed with FP64 atomics, so two of its runs may differ in the last
  bits);
- net momentum stays zero (pair forces are exact negatives, so only the
  rounding of the per-atom sums is left);
- the thermo rows (T, PE/N, E/N every 5 steps, over two list rebuilds)
  agree with the FP64 oracle's run of the same deck (1e-5 mixed, 1e-9
  FP64);
- forces on the final state (after the ghost fold for the half list) are
  within 1e-5 (mixed) or 1e-10 (FP64) relative L2 of the FP64 oracle's
  full-list forces on the same positions (cabanamd_impl.h:
  228-300 run loop, force_lj_cabana_neigh_impl.h:151-203).
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

NC = (80, 80, 80)
STEPS = 45


# precision, half list, relative tolerance of thermo rows and of forces
CASES = {"mixed-full": ("mixed", False, 1e-5, 1e-5),
         "fp64-full": ("fp64", False, 1e-9, 1e-10),
         "fp64-half": ("fp64", True, 1e-9, 1e-10)}


def _run(M, prec, half, fuse):
    p = M.lattice_params(ncells=NC, thermo=5, nsteps=STEPS, every=20, half=half)
    P = M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED
    md = M.CbnMD(p, precision=P, lanes=2, tile=6, fuse_ghosts=fuse)
    md.init()
    md.run(STEPS)
    return md


@pytest.mark.parametrize("case", list(CASES))
def test_lj_2M_full_size_properties(gpu, case):
    M = gpu
    prec, half, tol_th, tol_f = CASES[case]
    a = _run(M, prec, half, True)
    n = a.n_local
    assert n == 4 * NC[0] * NC[1] * NC[2]
    print(f"2M: layout {a.neighbor.layout()}", flush=True)
    xa, va, fa, ta, ia = a.arrays()
    tha = a.thermo()
    if not half:
        b = _run(M, prec, half, False)
        xb, vb, fb, tb, ib = b.arrays()
        assert a.neighbor.layout() == b.neighbor.layout()
        for u, w in ((xa, xb), (va, vb), (fa, fb), (ta, tb), (ia, ib)):
            assert np.array_equal(u, w)
        assert np.array_equal(tha, b.thermo())
        del b

    # unit masses: net momentum of the local atoms
    p = va[:n].sum(axis=0)
    vrms = np.sqrt((va[:n] ** 2).sum(axis=1).mean())
    print(f"2M: |P|/(N vrms) = {np.abs(p).max() / (n * vrms):.3e}", flush=True)
    assert np.abs(p).max() < 1e-9 * n * vrms

    # thermo rows (step, T, PE/N, E/N) against the FP64 oracle's loop
    ref = O.MD(O.lj_params(ncells=NC, thermo=5, nsteps=STEPS, half=half))
    ref.init()
    ref.run(STEPS)
    tho = ref.thermo()
    del ref
    assert tha.shape[0] == tho.shape[0] == STEPS // 5 + 1
    dev = np.abs(tha[:, 1:4] - tho[:, 1:4]).max(axis=0) / np.abs(tho[:, 1:4]).max(axis=0)
    print(f"2M: thermo deviation vs FP64 oracle (T, PE/N, E/N) {dev}", flush=True)
    assert np.array_equal(tha[:, 0], tho[:, 0])
    assert (dev < tol_th).all()

    # forces on the final positions against the FP64 oracle's full list
    a.compute_forces()
    x, _, f, t, _ = a.arrays()
    off, idx = O.neigh_build(x, n, 2.8, half=False)
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(x, t, n, 1, lj1, lj2, cs, off, idx, False)[:n]
    err = np.linalg.norm(f[:n] - fo) / np.linalg.norm(fo)
    print(f"2M: force rel L2 vs FP64 oracle {err:.3e}", flush=True)
    assert err < tol_f
