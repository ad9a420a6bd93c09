"""Writes the small golden fixtures under tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

The oracle is pinned first (tests/test_oracle_cpu.py: published RNG and hash
vectors, the LAMMPS LJ benchmark energy, closed forms and the reference's
unit-test invariants); these files freeze its outputs so that both the oracle
and the HIP engine are checked against the same recorded numbers.

  lj256_full.npz, lj256_half.npz  256-atom fcc LJ (4^3 cells, rho 0.8442,
      T 1.44, seed 87287), 20 steps of dt 0.005, rebuild every 20; x, v, f
      ordered by atom id after 0 and 20 steps, thermo rows every 5 steps.
  neigh_tst.npz  tstNeighbor.hpp geometry (1000 atoms, last 200 ghosts,
      rc 2.32, box [-5.3 rc, 4.7 rc)), positions and brute-force full lists.
  rng.npz  loop-geom RNG: seeds after reset for 16 (ibase, coord) pairs and
      the next three uniforms of each.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import pyoracle as O  # noqa: E402


def lj256(half):
    p = O.lj_params((4, 4, 4), half=half, thermo=5, nsteps=20)
    md = O.MD(p)
    md.init()
    out = {}

    def grab(tag):
        x, v, f, t, ids = md.arrays()
        n = md.n_local
        o = np.argsort(ids[:n])
        out["x" + tag] = x[:n][o]
        out["v" + tag] = v[:n][o]
        out["f" + tag] = f[:n][o]

    grab("0")
    md.run(20)
    grab("20")
    out["thermo"] = md.thermo()
    return out


def neigh_tst():
    rc = 2.32
    rng = np.random.default_rng(342343901)
    x = rng.uniform(-5.3 * rc, 4.7 * rc, (1000, 3))
    off, idx = O.neigh_brute(x, 800, rc, half=False)
    return dict(x=x, n_local=np.int32(800), rc=np.float64(rc), offsets=off,
                idx=idx)


def rng():
    r = np.random.default_rng(11)
    ib = r.integers(1, 2**31 - 1, 16).astype(np.int32)
    co = r.uniform(-40, 40, (16, 3))
    seeds = np.zeros(16, np.int64)
    u = np.zeros((16, 3))
    for k in range(16):
        s = O.rng_reset(int(ib[k]), co[k])
        seeds[k] = s
        for m in range(3):
            u[k, m], s = O.rng_uniform(s)
    return dict(ibase=ib, coord=co, seed=seeds, uniform=u)


def main():
    np.savez_compressed(os.path.join(HERE, "lj256_full.npz"), **lj256(False))
    np.savez_compressed(os.path.join(HERE, "lj256_half.npz"), **lj256(True))
    np.savez_compressed(os.path.join(HERE, "neigh_tst.npz"), **neigh_tst())
    np.savez_compressed(os.path.join(HERE, "rng.npz"), **rng())


if __name__ == "__main__":
    main()
