"""The LJ timestep at its benchmark size: fcc rho 0.8442, 80^3 cells =
2,048,000 atoms, rc 2.5, skin 0.3, full list rebuilt every 20 steps,
dt 0.005, mixed precision, two lanes per atom on tiles six cells tall (the
layout the autotuner picks for this workload).  At this size
the checks are properties that do not depend on the size:
- the step with the ghost refresh fused into it equals the separate
  refresh pass bit for bit (positions, velocities, forces, thermo rows);
- net momentum stays zero (pair forces are exact negatives, so only the
  rounding of the per-atom sums is left);
- the thermo rows (T, PE/N, E/N every 5 steps, over two list rebuilds)
  agree with the FP64 oracle's run of the same deck within 1e-5;
- forces on the final state are within 1e-5 (relative L2) of the FP64
  oracle's full-list forces on the same positions (cabanamd_impl.h:
  228-300 run loop, force_lj_cabana_neigh_impl.h:151-203).
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

NC = (80, 80, 80)
STEPS = 45


def _run(M, fuse):
    p = M.lattice_params(ncells=NC, thermo=5, nsteps=STEPS, every=20)
    md = M.CbnMD(p, precision=M.PREC_MIXED, lanes=2, tile=6, fuse_ghosts=fuse)
    md.init()
    md.run(STEPS)
    return md


def test_lj_2M_full_size_properties(gpu):
    M = gpu
    a = _run(M, True)
    n = a.n_local
    assert n == 4 * NC[0] * NC[1] * NC[2]
    print(f"2M: layout {a.neighbor.layout()}", flush=True)
    xa, va, fa, ta, ia = a.arrays()
    tha = a.thermo()
    b = _run(M, False)
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
    ref = O.MD(O.lj_params(ncells=NC, thermo=5, nsteps=STEPS))
    ref.init()
    ref.run(STEPS)
    tho = ref.thermo()
    del ref
    assert tha.shape[0] == tho.shape[0] == STEPS // 5 + 1
    dev = np.abs(tha[:, 1:4] - tho[:, 1:4]).max(axis=0) / np.abs(tho[:, 1:4]).max(axis=0)
    print(f"2M: thermo deviation vs FP64 oracle (T, PE/N, E/N) {dev}", flush=True)
    assert np.array_equal(tha[:, 0], tho[:, 0])
    assert (dev < 1e-5).all()

    # forces on the final positions against the FP64 oracle's full list
    a.compute_forces()
    x, _, f, t, _ = a.arrays()
    off, idx = O.neigh_build(x, n, 2.8, half=False)
    lj1, lj2, cs = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    fo = O.lj_force(x, t, n, 1, lj1, lj2, cs, off, idx, False)[:n]
    err = np.linalg.norm(f[:n] - fo) / np.linalg.norm(fo)
    print(f"2M: force rel L2 vs FP64 oracle {err:.3e}", flush=True)
    assert err < 1e-5
