"""Decks end to end on the GPU: the reference's multi-region example
input/in.lb (four regions and types, per-type masses and velocities, only
the 1-1 pair interacting) run through CbnMD and through the cabanamd driver,
against the FP64 oracle's run of the same parsed deck (cabanamd_impl.h:
70-432); the driver's printed rows and final blocks field by field in the
reference's format (print_summary cabanamd_impl.h:647-668, the
PERFORMANCE / FRACTION block :400-421); --vacuum (inputCL.cpp:243-250).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import pyoracle as O
from test_deck_regions_cpu import IN_LB

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cabanamd_amd", "cabanamd")


def _deck(tmp_path, text=IN_LB):
    p = tmp_path / "in.lb"
    p.write_text(text)
    return p


@pytest.fixture(scope="module")
def oracle_lb(tmp_path_factory):
    import cabanamd_amd as M
    prm = M.InputFile().read_file(_deck(tmp_path_factory.mktemp("lb")))
    ref = O.MD(O.params_from(prm))
    ref.init()
    ref.run(100)
    return ref


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_in_lb_matches_oracle(gpu, tmp_path, oracle_lb, prec):
    M = gpu
    prm = M.InputFile().read_file(_deck(tmp_path))
    P = M.PREC_FP64 if prec == "fp64" else M.PREC_MIXED
    md = M.CbnMD(prm, precision=P, lanes=1 if prec == "fp64" else 0)
    md.init()
    assert md.n_local == 5000
    md.run(100)
    th, tho = md.thermo(), oracle_lb.thermo()
    assert th.shape[0] == tho.shape[0] == 11
    x, v, f, t, i = md.arrays()
    xo, vo, fo, to, io = oracle_lb.arrays()
    n = md.n_local
    oa, ob = np.argsort(i[:n]), np.argsort(io[:n])
    assert np.array_equal(t[:n][oa], to[:n][ob])
    print(f"{prec}: max |dE/N| {np.abs(th[:, 3] - tho[:, 3]).max():.3e}", flush=True)
    if prec == "fp64":
        assert np.allclose(th[:, :4], tho, rtol=1e-10, atol=1e-12)
        assert np.abs(x[:n][oa] - xo[:n][ob]).max() < 1e-9
        assert np.abs(v[:n][oa] - vo[:n][ob]).max() < 1e-9
    else:
        assert np.allclose(th[:, :4], tho, rtol=1e-5, atol=1e-6)


ROW = re.compile(r"^(\d+)\t(-?\d+\.\d{6})\t(-?\d+\.\d{6})\t(-?\d+\.\d{6})\t"
                 r"(\d+\.\d{2})\t(\d\.\d{2}e[+-]\d{2})$")


def test_driver_output_format(gpu, tmp_path, oracle_lb):
    r = subprocess.run([EXE, "-il", str(_deck(tmp_path)), "--lanes", "1"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    h = lines.index("#Timestep Temperature PotE ETot Time Atomsteps/s ")
    rows = []
    for ln in lines[h + 1:]:
        m = ROW.match(ln)
        if not m:
            break
        rows.append(m.groups())
    assert [int(g[0]) for g in rows] == list(range(0, 101, 10))
    assert rows[0][4] == "0.00" and rows[0][5] == "0.00e+00"
    tho = oracle_lb.thermo()
    for g, o in zip(rows, tho):
        for k in (1, 2, 3):
            assert abs(float(g[k]) - o[k]) <= 0.6e-6 + 1e-9 * abs(o[k]), (g, o)
    k = lines.index("#Procs Atoms | Time T_Force T_Neigh T_Comm T_Int T_lb T_Other |")
    perf = re.match(r"^1 5000 \| ((\d+\.\d{2} ){7})\| PERFORMANCE$", lines[k + 1])
    frac = re.match(r"^1 5000 \| 1\.00 ((\d+\.\d{2} ){6})\| FRACTION$", lines[k + 2])
    assert perf and frac, lines[k:k + 3]
    assert lines[k + 3] == ""
    assert lines[k + 4] == "#Steps/s Atomsteps/s Atomsteps/(proc*s)"
    sps = lines[k + 5].split()
    assert len(sps) == 3 and all(re.match(r"^\d\.\d{2}e[+-]\d{2}$", s) for s in sps)
    assert sps[1] == sps[2]


def test_driver_vacuum(gpu, tmp_path):
    deck = _deck(tmp_path)
    r = subprocess.run([EXE, "-il", str(deck), "--lanes", "1", "--vacuum", "1.5"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "1 5000 | " in r.stdout
