"""cmd_thermo_kinetic_noalloc (no reference counterpart; for callers that
cannot allocate device memory) against cmd_thermo_kinetic, whose Temperature
/ KinE sum it restates (property_kine_impl.h:55-76): the same kernels on the
same data give the same bits.  Two calls in a row, the second on a smaller
system, so a stale partial sum left in the shared scratch array would show;
then the same on a larger system again, and the capacity error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(M, nc, seed):
    p = M.lattice_params(ncells=nc, temp=1.0 + 0.1 * seed, seed=87287 + seed, gaussian=True)
    md = M.CbnMD(p, precision=M.PREC_FP64, lanes=1)
    md.init()
    return md


def _ke(M, md, fn):
    out = C.c_double()
    M._lib.call(fn, getattr(md, "_sys", None) or md.handle, C.byref(out))
    return out.value


def test_kinetic_noalloc_matches_allocating_entry_point(gpu):
    M = gpu
    big = _system(M, (12, 12, 12), 1)
    small = _system(M, (3, 4, 5), 2)
    for md in (big, small, big, small):
        ref = _ke(M, md, "cmd_thermo_kinetic")
        got = _ke(M, md, "cmd_thermo_kinetic_noalloc")
        print(f"n={md.n_local}: {got!r} vs {ref!r}", flush=True)
        assert got == ref
        # the FP64 oracle's sum over the downloaded velocities
        _, v, _, _, _ = md.arrays()
        n = md.n_local
        ke = 0.5 * float((v[:n] ** 2).sum())
        assert abs(got - ke) <= 1e-12 * ke


def test_kinetic_noalloc_after_a_run(gpu):
    """After MD steps (velocities changed in place by the fused step), the
    no-allocation sum still equals the allocating one."""
    M = gpu
    md = _system(M, (6, 6, 6), 3)
    md.run(25)
    assert _ke(M, md, "cmd_thermo_kinetic_noalloc") == _ke(M, md, "cmd_thermo_kinetic")


def test_kinetic_noalloc_rejects_more_than_its_capacity(gpu):
    """One atom past 32,768 blocks of 256: a clear capacity error (the
    allocating entry point takes the same system), and the next call on a
    smaller system is right again."""
    M = gpu
    n = 32768 * 256 + 1
    s = M.System(1)
    s.set_units(0.005)
    s.set_mass([2.0])
    v = np.random.default_rng(5).normal(size=(n, 3))
    s.upload(np.zeros((n, 3)), v)
    with pytest.raises(M.CmdError) as e:
        _ke(M, s, "cmd_thermo_kinetic_noalloc")
    assert e.value.status == M._lib.CMD_ERR_CAPACITY
    assert "8388608" in str(e.value)
    ke = _ke(M, s, "cmd_thermo_kinetic")
    assert abs(ke - float((v ** 2).sum())) <= 1e-12 * ke
    del s
    md = _system(M, (4, 4, 4), 4)
    assert _ke(M, md, "cmd_thermo_kinetic_noalloc") == _ke(M, md, "cmd_thermo_kinetic")
