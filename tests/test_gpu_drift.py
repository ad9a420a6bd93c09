"""NVE total-energy drift over 1000 steps on the reference's CPU-runnable LJ
case: fcc rho 0.8442, 20^3 cells = 32,000 atoms, rc 2.5, skin 0.3, full list
rebuilt every 20 steps, Gaussian velocities at T 1.44 (inputFile.h:94-107),
dt 0.005, thermo every 100 steps (cabanamd_impl.h:285-399 run loop,
:365-372 thermo rows).

- FP64 with one lane per atom sums every force in the oracle's order, so the
  1000-step trajectory and its thermo rows equal the FP64 oracle's.
- FP64 with the layout init tunes, and the mixed-precision step (FP32 pair
  factors), follow the oracle's trajectory only to rounding, which the
  chaotic dynamics amplify; their total energy per atom stays within 1e-5
  (FP64) and 1e-4 (mixed) of the oracle's at every thermo row, and their
  drift max |E - E0| / |E0| is within 2e-6 of the oracle's drift plus 30 %
  of it.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

NC = (20, 20, 20)
STEPS = 1000


def _drift(e):
    return float(np.abs(e - e[0]).max() / abs(e[0]))


@pytest.fixture(scope="module")
def oracle_rows():
    ref = O.MD(O.lj_params(ncells=NC, thermo=100, nsteps=STEPS, gaussian=True))
    ref.init()
    ref.run(STEPS)
    return ref.thermo()


@pytest.mark.parametrize("case", ["fp64-1lane", "fp64-tuned", "mixed-tuned"])
def test_c1_energy_drift_1000_steps(gpu, oracle_rows, case):
    M = gpu
    prec, layout = case.split("-")
    P = M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED
    p = M.lattice_params(ncells=NC, thermo=100, nsteps=STEPS, gaussian=True)
    md = M.CbnMD(p, precision=P, lanes=1 if layout == "1lane" else 0)
    md.init()
    assert md.n_local == 32000
    md.run(STEPS)
    th, tho = md.thermo(), oracle_rows
    assert th.shape[0] == tho.shape[0] == STEPS // 100 + 1
    assert np.array_equal(th[:, 0], tho[:, 0])
    d_gpu, d_orc = _drift(th[:, 3]), _drift(tho[:, 3])
    print(f"{case}: lanes {md.tune_info()[:2]}, drift {d_gpu:.4e} vs oracle "
          f"{d_orc:.4e}, max |dE/N| {np.abs(th[:, 3] - tho[:, 3]).max():.3e}",
          flush=True)
    if case == "fp64-1lane":
        assert np.allclose(th[:, :4], tho, rtol=1e-12, atol=1e-14)
        assert abs(d_gpu - d_orc) <= 1e-9
        return
    tol = 1e-5 if prec == "fp64" else 1e-4
    assert np.abs(th[:, 3] - tho[:, 3]).max() <= tol * np.abs(tho[:, 3]).max()
    assert abs(d_gpu - d_orc) <= 2e-6 + 0.3 * d_orc
