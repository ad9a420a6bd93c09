"""Multi-region decks (inputFile_impl.h:241-297 region / create_atoms,
:396-410 velocity per type, :755-777 typing lattice sites by region,
:795-865 per-type velocity seeds and temperatures) and --vacuum
(inputCL.cpp:243-250, inputFile_impl.h:553-559), on the host.

The deck below is the reference's load-balancing example input/in.lb: four
block regions (mid-box lies inside lo-box), four types with their own
masses, velocity seeds and temperatures.

Pins:
- the oracle's rand() is the C library's: its first values equal libc
  rand() after srand(1);
- every site is typed by the regions that hold it: sites only in lo-box are
  type 1, in hi-box 3, in hj-box 4; a site in lo-box and mid-box takes the
  type of one of the two, chosen by the next rand() value modulo 2 in one of
  the two region orders (the map's), one draw per created atom in lattice
  loop order;
- velocities follow an independent numpy restatement: per-type seeds through
  the oracle's loop-geom generator (pinned by its published vectors in
  test_oracle_cpu.py), zero net momentum, each atom scaled by
  sqrt(T_target(type) / T);
- the engine's parser and lattice equal the oracle's bit for bit.
"""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

import cabanamd_amd as M
import pyoracle as O

IN_LB = """# Example for exercising load balancing (not intended to be physical)

units		lj
atom_style	atomic

newton off

lattice		fcc 0.8442
region		lo-box block 0 10 0 10 0 10
region		mid-box block 5 10 5 10 5 10
region		hi-box block 0 5 15 20 15 20
region		hj-box block 15 20 15 20 0 5
create_box	4 box

mass		1 2.0
mass		2 8.0
mass		3 9.0
mass		4 1.0

create_atoms	1 region lo-box
create_atoms	2 region mid-box
create_atoms	3 region hi-box
create_atoms	4 region hj-box

velocity	1 create 10.4 87287 loop geom
velocity	2 create 41.6 87287 loop geom
velocity	3 create 20.8 87287 loop geom
velocity	4 create 10.4 87287 loop geom

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

NAMES = ["lo-box", "mid-box", "hi-box", "hj-box"]
BLOCKS = [(0, 10, 0, 10, 0, 10), (5, 10, 5, 10, 5, 10), (0, 5, 15, 20, 15, 20),
          (15, 20, 15, 20, 0, 5)]


def _parse(tmp_path, text=IN_LB):
    p = tmp_path / "in.lb"
    p.write_text(text)
    return M.InputFile().read_file(p)


def _domain(prm, vacuum=1.0):
    a = prm.lattice_constant
    lo = [a * prm.region_lo[k] for k in range(3)]
    hi = [a * prm.region_hi[k] * vacuum for k in range(3)]
    return M.create_domain(lo, hi, prm.comm_ghost_cutoff), \
        O.create_domain(lo, hi, prm.comm_ghost_cutoff)


def test_crand_is_the_c_library_rand():
    libc = C.CDLL(ctypes.util.find_library("c"))
    libc.srand(1)
    want = [libc.rand() for _ in range(5000)]
    assert list(O.crand(5000)) == want


def test_in_lb_parses(tmp_path):
    prm = _parse(tmp_path)
    assert prm.nregions == 4 and prm.ntypes == 4
    assert [prm.region_name[k].value.decode() for k in range(4)] == NAMES
    for k, b in enumerate(BLOCKS):
        assert list(prm.region_block[k]) == [float(v) for v in b]
    assert list(prm.region_type)[:4] == [1, 2, 3, 4]
    assert list(prm.region_lo) == [0.0, 0.0, 0.0]
    assert list(prm.region_hi) == [20.0, 20.0, 20.0]
    assert list(prm.mass)[:4] == [2.0, 8.0, 9.0, 1.0]
    assert prm.temp_target == 10.4 and prm.temp_seed == 87287
    assert list(prm.type_temp)[1:4] == [41.6, 20.8, 10.4]
    assert list(prm.type_seed)[1:4] == [87287] * 3
    assert prm.ncoeff == 1 and prm.coeff_cut[0] == 2.5
    assert prm.vacuum_rate == 1.0


@pytest.mark.parametrize("line,msg", [
    ("create_atoms 1 region nowhere", "region 'nowhere' is not defined"),
    ("region r sphere 0 0 0 1", "'region' command only supports 'block'"),
    ("velocity 2 set 1.0 5", "can only be used with option 'create'"),
])
def test_region_deck_errors(tmp_path, line, msg):
    with pytest.raises(M.CmdError) as e:
        _parse(tmp_path, "units lj\nlattice fcc 0.8442\n"
                         "region box block 0 2 0 2 0 2\n" + line + "\n")
    assert msg in str(e.value)


def test_create_atoms_box_names_a_region(tmp_path):
    """create_atoms T box types the regions map's first entry; a deck with
    one region (in.lj) makes every site type T."""
    prm = _parse(tmp_path, "units lj\nlattice fcc 0.8442\n"
                           "region box block 0 3 0 3 0 3\ncreate_box 2 box\n"
                           "create_atoms 2 box\nvelocity 2 create 1.0 5\n")
    assert prm.nregions == 1 and prm.region_type[0] == 2
    d, _ = _domain(prm)
    inp = M.InputFile()
    inp.params = prm
    x, v, t, i = inp.create_lattice(d)
    assert len(x) == 4 * 27 and np.all(t == 1)


def _region_hits(x, a):
    hits = []
    for k, b in enumerate(BLOCKS):
        lo = a * np.array(b[0::2], float)
        hi = a * np.array(b[1::2], float)
        hits.append(np.all((x >= lo) & (x < hi), axis=1))
    return np.array(hits).T


def test_in_lb_lattice_types_by_region(tmp_path):
    prm = _parse(tmp_path)
    d, do = _domain(prm)
    inp = M.InputFile()
    inp.params = prm
    x, v, t, i = inp.create_lattice(d)
    a = prm.lattice_constant
    h = _region_hits(x, a)
    # 10^3 + 5^3 + 5^3 unit cells, 4 sites each (mid-box inside lo-box)
    assert len(x) == 4 * (1000 + 125 + 125)
    assert np.array_equal(i, np.arange(1, len(x) + 1))
    assert np.all(h.any(axis=1))
    only_lo = h[:, 0] & ~h[:, 1]
    assert np.all(t[only_lo] == 0)
    assert np.all(t[h[:, 2]] == 2) and np.all(t[h[:, 3]] == 3)
    both = h[:, 0] & h[:, 1]
    assert both.sum() == 500
    # one rand() per created atom, in creation order; mod 2 picks among
    # (lo-box, mid-box) in one of the two orders
    r = O.crand(len(x))[both] % 2
    order_a = np.where(r == 0, 0, 1)
    order_b = np.where(r == 0, 1, 0)
    assert np.array_equal(t[both], order_a) or np.array_equal(t[both], order_b)
    assert 150 < (t[both] == 1).sum() < 350


def test_in_lb_velocities_independent_restatement(tmp_path):
    prm = _parse(tmp_path)
    d, _ = _domain(prm)
    inp = M.InputFile()
    inp.params = prm
    x, v, t, _ = inp.create_lattice(d)
    temps = [prm.temp_target] + list(prm.type_temp)[1:4]
    seeds = [prm.temp_seed] + list(prm.type_seed)[1:4]
    mass = np.array(list(prm.mass)[:4])
    u = np.zeros_like(x)
    for k in range(len(x)):
        s = O.rng_reset(seeds[t[k]], x[k])
        for c in range(3):
            val, s = O.rng_uniform(s)
            u[k, c] = val - 0.5
    m = mass[t]
    u /= np.sqrt(m)[:, None]
    u -= (m[:, None] * u).sum(axis=0) / m.sum()
    T = (m[:, None] * u * u).sum() / (3 * len(x) - 3)
    u *= np.sqrt(np.array(temps)[t] / T)[:, None]
    assert np.allclose(v, u, rtol=1e-12, atol=1e-15)


def test_in_lb_lattice_matches_oracle(tmp_path):
    prm = _parse(tmp_path)
    d, do = _domain(prm)
    inp = M.InputFile()
    inp.params = prm
    x, v, t, i = inp.create_lattice(d)
    po = O.params_from(prm)
    xo, to, io = O.lattice(po, do)
    vo = O.velocity_create(po, xo, to)
    assert np.array_equal(x, xo) and np.array_equal(t, to)
    assert np.array_equal(i, io) and np.array_equal(v, vo)


def test_vacuum_grows_the_box_not_the_lattice(tmp_path):
    prm = _parse(tmp_path)
    d1, _ = _domain(prm)
    d2, _ = _domain(prm, vacuum=1.5)
    inp = M.InputFile()
    inp.params = prm
    x1, _, t1, _ = inp.create_lattice(d1)
    x2, _, t2, _ = inp.create_lattice(d2)
    assert list(d2.high) == [1.5 * h for h in d1.high]
    assert np.array_equal(x1, x2) and np.array_equal(t1, t2)
