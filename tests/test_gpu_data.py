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


def test_dump_and_correctness(gpu, tmp_path):
    """Snapshots of one run checked against a second identical run (all
    differences zero) and against a mixed-precision run (small)."""
    M = gpu
    ref = os.path.join(tmp_path, "ref")
    os.makedirs(ref)
    p = M.lattice_params(ncells=(5, 5, 5), thermo=10, nsteps=20)
    a = M.CbnMD(p, precision=M.PREC_FP64, lanes=1)
    a.init()
    a.dump(ref, 0)
    a.run(10)
    a.dump(ref, 10)
    assert os.path.exists(os.path.join(ref, "output.0000000010.000"))
    log = os.path.join(tmp_path, "correctness.dat")
    b = M.CbnMD(p, precision=M.PREC_FP64, lanes=1)
    b.init()
    assert np.array_equal(b.check_correctness(ref, 0, log), np.zeros(6))
    b.run(10)
    assert np.array_equal(b.check_correctness(ref, 10, log), np.zeros(6))
    lines = open(log).read().splitlines()
    assert lines[0].startswith("# timestep deltarnorm")
    assert [ln.split()[0] for ln in lines[1:]] == ["0", "10"]
    c = M.CbnMD(p, precision=M.PREC_MIXED, lanes=2)
    c.init()
    c.run(10)
    d = c.check_correctness(ref, 10)
    assert 0 < d[1] < 1e-6 and d[5] < 1e-3
    with pytest.raises(M.CmdError):
        c.check_correctness(ref, 20)  # no such snapshot


def test_cli_dumpbinary_and_correctness(gpu, tmp_path):
    import subprocess
    import sys
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "cabanamd_amd", "cabanamd")
    deck = os.path.join(tmp_path, "in.lj")
    with open(deck, "w") as fh:
        fh.write("units lj\natom_style atomic\nlattice fcc 0.8442\n"
                 "region box block 0 5 0 5 0 5\ncreate_box 1 box\n"
                 "create_atoms 1 box\nmass 1 1.0\nvelocity all create 1.44 87287 "
                 "loop geom\npair_style lj/cut 2.5\npair_coeff 1 1 1.0 1.0 2.5\n"
                 "neighbor 0.3 bin\nneigh_modify every 20 one 50\n"
                 "fix 1 all nve\nthermo 10\nrun 20\n")
    ref = os.path.join(tmp_path, "snap")
    os.makedirs(ref)
    r = subprocess.run([exe, "-il", deck, "--lanes", "1", "--dumpbinary", "5", ref],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert len([f for f in os.listdir(ref) if f.startswith("output.")]) == 5
    out = os.path.join(tmp_path, "corr.dat")
    r = subprocess.run([exe, "-il", deck, "--lanes", "1", "--correctness", "10", ref, out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [ln.split() for ln in open(out).read().splitlines()[1:]]
    assert [row[0] for row in rows] == ["0", "10", "20"]
    assert all(float(v) == 0.0 for row in rows for v in row[1:])


def test_kernel_profile_every_kth_launch(gpu):
    """cmd_md_profile(k > 1) times one fused force launch in k and no other
    phase; k = 1 times every phase of every step."""
    M = gpu
    p = M.lattice_params(ncells=(8, 8, 8), thermo=0, nsteps=40, gaussian=True)
    md = M.CbnMD(p, precision=M.PREC_MIXED, lanes=2)
    md.init()
    md.profile(4)
    md.run(20)
    n4, ms4 = md.kernel_stats()
    t4 = md.timers()
    assert n4 == 5 and ms4 > 0
    assert t4["neigh"] == 0 and t4["comm"] == 0 and t4["force"] > 0
    md.profile(1)
    md.run(20)
    n1, ms1 = md.kernel_stats()
    t1 = md.timers()
    assert n1 == 20 and ms1 > 0
    assert t1["neigh"] > 0  # the rebuild at step 40
    md.profile(0)
