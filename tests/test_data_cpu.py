"""LAMMPS data files (read_data.h:100-422): the format the reference reads
and writes either side of the timestep path.  No GPU needed."""
import os

import numpy as np
import pytest

import cabanamd_amd as M

DATA = """LAMMPS data file for a test

4 atoms
2 atom types

0.0 10.0 xlo xhi
-1.0 9.0 ylo yhi
0.5 10.5 zlo zhi

Masses

1 1.5
2 4.0

Atoms # atomic

3 2 1.25 2.5 3.75
1 1 0.0 -1.0 0.5   # on the low faces: kept
4 1 9.5 8.5 10.0
2 2 5.0 5.0 5.0

Velocities

2 0.2 0.0 -0.2
1 0.1 0.2 0.3
4 -1 -2 -3
3 3.5 0 0
"""


def _write(tmp_path, text, name="in.data"):
    p = os.path.join(tmp_path, name)
    with open(p, "w") as fh:
        fh.write(text)
    return p


def test_read_data_sections(tmp_path):
    d = M.read_data(_write(tmp_path, DATA))
    assert d["ntypes"] == 2 and list(d["mass"]) == [1.5, 4.0]
    assert list(d["lo"]) == [0.0, -1.0, 0.5] and list(d["hi"]) == [10.0, 9.0, 10.5]
    assert list(d["id"]) == [3, 1, 4, 2]
    assert list(d["type"]) == [1, 0, 0, 1]  # 1-based in the file
    assert np.array_equal(d["x"][1], [0.0, -1.0, 0.5])
    # velocities are matched to atoms by id
    assert np.array_equal(d["v"][0], [3.5, 0, 0])
    assert np.array_equal(d["v"][1], [0.1, 0.2, 0.3])
    assert np.array_equal(d["v"][3], [0.2, 0.0, -0.2])


def test_read_data_charge_style(tmp_path):
    text = DATA.replace("3 2 1.25 2.5 3.75", "3 2 -0.5 1.25 2.5 3.75") \
               .replace("1 1 0.0 -1.0 0.5", "1 1 0.5 0.0 -1.0 0.5") \
               .replace("4 1 9.5 8.5 10.0", "4 1 0 9.5 8.5 10.0") \
               .replace("2 2 5.0 5.0 5.0", "2 2 0 5.0 5.0 5.0")
    d = M.read_data(_write(tmp_path, text), charge=True)
    assert list(d["q"]) == [-0.5, 0.5, 0.0, 0.0]
    assert np.array_equal(d["x"][0], [1.25, 2.5, 3.75])


@pytest.mark.parametrize("bad, msg", [
    (DATA.replace("Masses", "Bonds"), "Unknown identifier"),
    (DATA.replace("4 1 9.5 8.5 10.0", "4 1 9.5 8.5 10.5"), "incorrect # of atoms"),
    (DATA[:DATA.index("Masses")] + DATA[DATA.index("Velocities"):], "Atoms before"),
])
def test_read_data_errors(tmp_path, bad, msg):
    with pytest.raises(M.CmdError) as e:
        M.read_data(_write(tmp_path, bad))
    assert msg in str(e.value)


def test_write_then_read_is_exact(tmp_path):
    x, v, t, i, lo, hi = M.fcc_system((3, 4, 5), 0.8, 1.3, 11)
    t = (i % 2).astype(np.int32)
    p = os.path.join(tmp_path, "out.data")
    M.write_data(p, x, v, t, i, lo, hi, 2)
    d = M.read_data(p)
    assert np.array_equal(d["x"], x) and np.array_equal(d["v"], v)
    assert np.array_equal(d["type"], t) and np.array_equal(d["id"], i)
    assert np.array_equal(d["hi"], hi)
    text = open(p).read()
    assert "Atoms # atomic" in text and "\nVelocities\n" in text


def test_write_data_wraps_positions(tmp_path):
    x = np.array([[-0.25, 5.0, 10.0], [10.5, -1e-18, 3.0]])
    p = os.path.join(tmp_path, "w.data")
    M.write_data(p, x, np.zeros_like(x), [0, 0], [1, 2], [0, 0, 0], [10, 10, 10], 1)
    d = M.read_data(p)
    assert np.allclose(d["x"], [[9.75, 5.0, 0.0], [0.5, 0.0, 3.0]], atol=1e-15)


def test_deck_read_and_write_data_keywords(tmp_path):
    deck = _write(tmp_path, "units lj\natom_style atomic\nread_data in.data\n"
                  "write_data out.data\npair_style lj/cut 2.5\n"
                  "pair_coeff 1 1 1.0 1.0 2.5\nrun 10\n", "in.deck")
    inp = M.InputFile()
    p = inp.read_file(deck)
    assert p.data_in == b"in.data" and p.data_out == b"out.data"
    assert p.atom_style_charge == 0
