"""C-ABI library checks that need no GPU: it loads, it exports every entry
point include/cabanamd_amd.h declares, and its host-side logic (deck parser,
domain, lattice and velocity generation) matches the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cabanamd_amd as M
from cabanamd_amd import _lib
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(cmd_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.exported_names()) <= declared | {"cmd_last_error"}


def test_no_device_reports_cleanly():
    if M.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(M.CmdError) as e:
        M.System(1)
    assert e.value.status == _lib.CMD_ERR_NO_DEVICE


@pytest.mark.parametrize("gauss", [False, True])
@pytest.mark.parametrize("ncells", [(4, 4, 4), (6, 7, 8), (20, 20, 20)])
def test_lattice_and_velocities_match_oracle(ncells, gauss):
    pp = M.lattice_params(ncells=ncells, gaussian=gauss)
    po = O.lj_params(ncells=ncells, gaussian=gauss)
    a = pp.lattice_constant
    hi = [a * c for c in ncells]
    d = M.create_domain([0, 0, 0], hi, pp.comm_ghost_cutoff)
    do = O.create_domain([0, 0, 0], hi, po.comm_ghost_cutoff)
    for f in ["low", "high", "L", "cell", "local_lo", "local_hi",
              "local_mesh", "ghost_lo", "ghost_hi"]:
        assert list(getattr(d, f)) == list(getattr(do, f)), f
    assert d.halo_width == do.halo_width
    inp = M.InputFile()
    inp.params = pp
    x, v, t, i = inp.create_lattice(d)
    xo, to, io = O.lattice(po, do)
    vo = O.velocity_create(po, xo, to)
    assert x.shape[0] == 4 * np.prod(ncells)
    assert np.array_equal(x, xo)
    assert np.array_equal(v, vo)
    assert np.array_equal(i, io) and np.array_equal(t, to)


DECK = """# LJ melt, fcc 0.8442
units		lj
atom_style	atomic

newton off
lattice		fcc 0.8442
region		box block 0 40 0 40 0 40
create_box	1 box
create_atoms	1 box
mass		1 2.0

velocity	all create 1.4 87287 loop geom

pair_style	lj/cut 2.5
pair_coeff	1 1 1.0 1.0 2.5

neighbor	0.3 bin
neigh_modify	every 20 one 50
comm_modify	cutoff * 20
fix		1 all nve
thermo 10

dump dmpvtk all vtk 10 dump%_*.vtu

run		100
"""


def test_deck_parser(tmp_path):
    p = tmp_path / "in.lj"
    p.write_text(DECK)
    prm = M.InputFile().read_file(p)
    assert prm.lattice_constant == (4.0 / 0.8442) ** (1.0 / 3.0)
    assert list(prm.region_hi) == [40.0, 40.0, 40.0]
    assert prm.mass[0] == 2.0 and prm.temp_target == 1.4
    assert prm.temp_seed == 87287 and prm.force_cutoff == 2.5
    assert prm.skin == 0.3 and prm.every == 20 and prm.max_neigh_guess == 50
    assert prm.comm_ghost_cutoff == 20.0 and prm.half == 0
    assert prm.nsteps == 100 and prm.thermo_rate == 10 and prm.dt == 0.005
    assert prm.ncoeff == 1 and prm.coeff_cut[0] == 2.5
    # --force-iteration NEIGH_HALF overrides 'newton off'
    assert M.InputFile(force_iteration=1).read_file(p).half == 1
    p.write_text(DECK.replace("newton off", "newton on"))
    assert M.InputFile().read_file(p).half == 1


@pytest.mark.parametrize("line,msg", [
    ("variable a equal 1", "'variable' keyword is not supported"),
    ("units si", "'units' command only supports"),
    ("lattice bcc 1.0", "'lattice' command only supports 'sc' and 'fcc'"),
    ("fix 1 all nvt", "'fix' command only supports 'nve'"),
    ("newton maybe", "'newton' must be followed by 'on' or 'off'"),
    ("frobnicate 3", "Unknown input file keyword"),
])
def test_deck_errors_like_reference(tmp_path, line, msg):
    p = tmp_path / "in.bad"
    p.write_text("units lj\n" + line + "\n")
    with pytest.raises(M.CmdError) as e:
        M.InputFile().read_file(p)
    assert msg in str(e.value)


def test_metal_units_force_timestep(tmp_path):
    p = tmp_path / "in.m"
    p.write_text("timestep 0.002\nunits metal\nlattice fcc 3.52\n"
                 "region box block 0 3 0 3 0 3\n")
    prm = M.InputFile().read_file(p)
    # units metal sets dt = 0.001 regardless of an earlier timestep
    assert prm.dt == 0.001 and prm.lattice_constant == 3.52
    assert prm.boltz == 8.617343e-5
