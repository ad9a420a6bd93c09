"""Pins the FP64 oracle (oracle/liboracle.so) against facts fixed outside it.

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign, index or comparison, an unsigned/signed mix-up) fails it:
published known answers (Park-Miller, Jenkins one-at-a-time, the LAMMPS LJ
benchmark energy), closed forms (velocity Verlet under constant force, the LJ
minimum and cutoff, F = -dE/dx), the reference's own unit tests restated
(tstNeighbor.hpp brute force and half-list invariants, tstIntegrator.hpp
reversibility) and invariants (momentum, permutation bijection, ghost images).
No GPU is used.
"""
import itertools
import math

import numpy as np
import pytest

import pyoracle as O

IM = 2147483647


# ---------------------------------------------------------------- RNG -------
def test_park_miller_minimal_standard_known_answer():
    """Park & Miller (CACM 1988): from seed 1 the 10000th state is 1043618065.
    Restated generator: inputFile.h:84-92."""
    s = 1
    for _ in range(10000):
        u, s = O.rng_uniform(s)
    assert s == 1043618065
    assert u == s * (1.0 / IM)


def _oaat(data, signed_chars=True):
    """Jenkins one-at-a-time hash (published algorithm); with signed_chars the
    byte is sign-extended before the add, as a `char` is on x86."""
    h = 0
    for b in data:
        if signed_chars and b >= 128:
            b -= 256
        h = (h + b) & 0xFFFFFFFF
        h = (h + (h << 10)) & 0xFFFFFFFF
        h ^= h >> 6
    h = (h + (h << 3)) & 0xFFFFFFFF
    h ^= h >> 11
    h = (h + (h << 15)) & 0xFFFFFFFF
    return h


def test_oaat_published_vectors():
    assert _oaat(b"a") == 0xCA2E9442
    assert _oaat(b"The quick brown fox jumps over the lazy dog") == 0x519E91F5


def _reset_expected(ibase, coord, signed_chars=True):
    """inputFile.h:108-147: hash the int then the three doubles (native
    bytes), keep 27 bits, never 0, warm up with five draws."""
    data = np.int32(ibase).tobytes() + np.asarray(coord, np.float64).tobytes()
    seed = _oaat(data, signed_chars) & 0x7FFFFFF
    seed = seed or 1
    for _ in range(5):
        _, seed = O.rng_uniform(seed)
    return seed


def test_rng_reset_matches_hash_with_signed_chars():
    rng = np.random.default_rng(5)
    differs = 0
    for k in range(200):
        ibase = int(rng.integers(1, 2**31 - 1))
        coord = rng.uniform(-50, 50, 3)
        got = O.rng_reset(ibase, coord)
        assert got == _reset_expected(ibase, coord)
        differs += got != _reset_expected(ibase, coord, signed_chars=False)
    # doubles carry bytes >= 0x80, so the sign extension matters
    assert differs > 150


def test_gaussian_moments():
    """Marsaglia polar method (inputFile.h:94-106): N(0,1) moments."""
    s = 12345
    g = np.empty(100000)
    for k in range(g.size):
        g[k], s = O.rng_gaussian(s)
    assert abs(g.mean()) < 0.015
    assert abs(g.var() - 1.0) < 0.02
    assert abs(np.mean(g ** 4) - 3.0) < 0.1


# ------------------------------------------------------------- lattice ------
def _setup(ncells=(4, 4, 4), **kw):
    p = O.lj_params(ncells, **kw)
    lo = np.array(p.region_lo[:]) * p.lattice_constant
    hi = np.array(p.region_hi[:]) * p.lattice_constant
    d = O.create_domain(lo, hi, p.comm_ghost_cutoff)
    return p, d


def test_fcc_lattice_geometry():
    p, d = _setup((5, 4, 3))
    x, t, ids = O.lattice(p, d)
    a = p.lattice_constant
    assert len(x) == 4 * 5 * 4 * 3
    assert sorted(ids) == list(range(1, len(x) + 1))
    assert np.all(t == 0)
    # every site is a*(i + basis), basis in {0, 1/2}^3 with even parity
    f = x / a * 2.0
    assert np.allclose(f, np.round(f), atol=1e-9)
    k = np.round(f).astype(int)
    assert np.all(k.sum(axis=1) % 2 == 0)
    # nearest-neighbour distance a / sqrt(2) (minimum image)
    L = np.array([5, 4, 3]) * a
    dx = x[:, None, :] - x[None, :, :]
    dx -= L * np.round(dx / L)
    r = np.sqrt((dx ** 2).sum(-1)) + np.eye(len(x)) * 1e9
    assert np.isclose(r.min(), a / math.sqrt(2), rtol=1e-12)
    assert np.all(np.sum(np.isclose(r, a / math.sqrt(2)), axis=1) == 12)


def test_velocity_zero_momentum_and_target_temperature():
    p, d = _setup((4, 4, 4), temp=1.44)
    x, t, ids = O.lattice(p, d)
    v = O.velocity_create(p, x, t)
    assert np.abs(v.sum(axis=0)).max() < 1e-13
    T = O.temperature(v, t, np.array([1.0]))
    assert abs(T - 1.44) < 1e-12
    # no two atoms share a velocity: each was drawn from its own seed
    assert len(np.unique(v[:, 0])) == len(v)


# ---------------------------------------------------------- integrator ------
def test_velocity_verlet_constant_force_closed_form():
    """x(n) = x0 + n dt v0 + n^2 dt^2 a / 2, v(n) = v0 + n dt a."""
    rng = np.random.default_rng(1)
    n = 50
    x0 = rng.uniform(-1, 1, (n, 3))
    v0 = rng.uniform(-1, 1, (n, 3))
    f = rng.uniform(-1, 1, (n, 3))
    typ = rng.integers(0, 2, n).astype(np.int32)
    mass = np.array([1.0, 2.5])
    dt, mvv2e = 0.005, 1.0
    x, v = x0.copy(), v0.copy()
    steps = 40
    for _ in range(steps):
        O.nve_initial(x, v, f, typ, mass, dt, mvv2e)
        O.nve_final(v, f, typ, mass, dt, mvv2e)
    a = f / (mass[typ][:, None] * mvv2e)
    xe = x0 + steps * dt * v0 + 0.5 * (steps * dt) ** 2 * a
    ve = v0 + steps * dt * a
    assert np.allclose(x, xe, rtol=0, atol=1e-13)
    assert np.allclose(v, ve, rtol=0, atol=1e-13)


def test_integrator_reversibility():
    """tstIntegrator.hpp:83-137: 100 steps, negate v, 100 steps: back home."""
    rng = np.random.default_rng(2)
    n = 100
    x0 = rng.uniform(-3, 3, (n, 3))
    v = rng.uniform(-1, 1, (n, 3))
    f = np.zeros((n, 3))
    typ = np.zeros(n, np.int32)
    mass = np.array([1.0])
    x = x0.copy()
    for _ in range(100):
        O.nve_initial(x, v, f, typ, mass, 0.005)
        O.nve_final(v, f, typ, mass, 0.005)
    v = -v
    for _ in range(100):
        O.nve_initial(x, v, f, typ, mass, 0.005)
        O.nve_final(v, f, typ, mass, 0.005)
    ulp = np.spacing(np.abs(x0).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(x - x0) <= 4 * ulp)


# ------------------------------------------------------------------ LJ ------
def _pair_force_energy(r, rc=2.5):
    x = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    t = np.zeros(2, np.int32)
    lj1, lj2, cutsq = O.lj_coeff(1, [(1, 1, 1.0, 1.0, rc)])
    off, idx = O.neigh_brute(x, 2, rc + 0.3, half=False)
    f = O.lj_force(x, t, 2, 1, lj1, lj2, cutsq, off, idx, False)
    e = O.lj_energy(x, t, 2, 1, lj1, lj2, cutsq, off, idx, False)
    return f, e


def test_lj_minimum_and_cutoff():
    f, e = _pair_force_energy(2.0 ** (1.0 / 6.0))
    assert np.abs(f).max() < 1e-13
    # shifted pair energy at the minimum: -eps - E_lj(rc)
    erc = 4.0 * (2.5 ** -12 - 2.5 ** -6)
    assert abs(e - (-1.0 - erc)) < 1e-13
    # shifted energy goes to zero at the cutoff; beyond it nothing acts
    f, e = _pair_force_energy(2.5 * (1 - 1e-12))
    assert abs(e) < 1e-11
    f, e = _pair_force_energy(2.5 * (1 + 1e-12))
    assert e == 0.0 and np.all(f == 0.0)


def test_lj_force_is_minus_energy_gradient():
    """Central difference of the (shifted) energy: the shift is a constant
    inside rc, so F = -dE/dx holds for every atom."""
    rng = np.random.default_rng(3)
    base = np.array(list(itertools.product(range(3), repeat=3)), float) * 1.15
    x = base + rng.uniform(-0.08, 0.08, base.shape)
    n = len(x)
    t = (np.arange(n) % 2).astype(np.int32)
    lj1, lj2, cutsq = O.lj_coeff(2, [(1, 1, 1.0, 1.0, 2.5), (1, 2, 0.8, 1.1, 2.5),
                                     (2, 2, 1.3, 0.9, 2.5)])
    off, idx = O.neigh_brute(x, n, 2.8, half=False)
    f = O.lj_force(x, t, n, 2, lj1, lj2, cutsq, off, idx, False)
    assert np.abs(f.sum(axis=0)).max() < 1e-10  # Newton's third law
    h = 1e-6
    for i in (0, 5, 13, 26):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            ep = O.lj_energy(xp, t, n, 2, lj1, lj2, cutsq, off, idx, False)
            em = O.lj_energy(xm, t, n, 2, lj1, lj2, cutsq, off, idx, False)
            # the energy sums each pair from both sides with weight 1/2
            fd = -(ep - em) / (2 * h)
            assert abs(fd - f[i, k]) < 1e-6 * max(1.0, abs(f[i, k]))


def test_lj_benchmark_lattice_energy():
    """LAMMPS bench/in.lj (fcc, rho 0.8442, rc 2.5, unshifted LJ) prints
    E_pair = -6.7733681 at step 0.  The oracle reports the shifted energy
    (force_lj_cabana_neigh_impl.h:261-315); adding back the shift times the
    pairs inside rc must give the published figure."""
    p, d = _setup((8, 8, 8))
    md = O.MD(p)
    md.init()
    rows = md.thermo()
    pe_shifted = rows[0][2]
    x, v, f, t, ids = md.arrays()
    n = md.n_local
    a = p.lattice_constant
    L = 8 * a
    xl = x[:n]
    dx = xl[:, None, :] - xl[None, :, :]
    dx -= L * np.round(dx / L)
    rsq = (dx ** 2).sum(-1)
    inside = (rsq < 2.5 ** 2) & (rsq > 0)
    pairs_per_atom = inside.sum() / n / 2.0
    erc = 4.0 * (2.5 ** -12 - 2.5 ** -6)
    assert abs(pe_shifted + pairs_per_atom * erc - (-6.7733681)) < 5e-8


def _lj_pair_virial_closed_form(r, eps, sigma, rc):
    """r . F(r) of one LJ pair: r (-dE/dr) = 48 eps (s/r)^12 - 24 eps (s/r)^6
    inside the cutoff (strict), 0 outside."""
    if r >= rc:
        return 0.0
    s6 = (sigma / r) ** 6
    return 48.0 * eps * s6 * s6 - 24.0 * eps * s6


@pytest.mark.parametrize("half", [False, True])
def test_lj_virial_two_atom_closed_form(half):
    """orc_lj_virial = sum over pairs of r_ij . F_ij: a single pair at
    distances on both sides of the minimum (sign change at 2^(1/6) sigma),
    next to the cutoff and beyond it; the full list sums each pair from
    both sides with weight 1/2, the half list once."""
    lj1, lj2, cutsq = O.lj_coeff(1, [(1, 1, 1.0, 1.0, 2.5)])
    t = np.zeros(2, np.int32)
    for r in (0.93, 1.05, 2.0 ** (1.0 / 6.0), 1.37, 2.1, 2.4999, 2.7):
        d = np.array([0.6, -0.48, 0.64])  # a unit direction off the axes
        x = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0]]) + np.outer([0, r], d)
        off, idx = O.neigh_brute(x, 2, 2.8, half=half)
        w = O.lj_virial(x, t, 2, 1, lj1, lj2, cutsq, off, idx, half)
        ref = _lj_pair_virial_closed_form(r, 1.0, 1.0, 2.5)
        assert abs(w - ref) <= 1e-12 * max(1.0, abs(ref)), (r, w, ref)
    # repulsive inside the minimum, attractive outside it
    assert _lj_pair_virial_closed_form(1.0, 1.0, 1.0, 2.5) > 0
    assert _lj_pair_virial_closed_form(1.5, 1.0, 1.0, 2.5) < 0


def test_lj_virial_mixed_types_sums_pairs():
    """Three atoms of two types with per-pair eps, sigma and cutoff: W is the
    sum of the three pairs' closed forms (a transposed type index or a
    dropped pair changes it)."""
    coeffs = [(1, 1, 1.0, 1.0, 2.5), (1, 2, 0.7, 1.2, 2.2), (2, 2, 1.4, 0.9, 2.0)]
    lj1, lj2, cutsq = O.lj_coeff(2, coeffs)
    x = np.array([[0.0, 0.0, 0.0], [1.1, 0.2, -0.1], [0.3, 1.25, 0.4]])
    t = np.array([0, 1, 1], np.int32)
    table = {(0, 0): coeffs[0], (0, 1): coeffs[1], (1, 0): coeffs[1],
             (1, 1): coeffs[2]}
    ref = 0.0
    for i, j in ((0, 1), (0, 2), (1, 2)):
        _, _, eps, sig, rc = table[(t[i], t[j])]
        ref += _lj_pair_virial_closed_form(np.linalg.norm(x[i] - x[j]), eps,
                                           sig, rc)
    for half in (False, True):
        off, idx = O.neigh_brute(x, 3, 2.8, half=half)
        w = O.lj_virial(x, t, 3, 2, lj1, lj2, cutsq, off, idx, half)
        assert abs(w - ref) <= 1e-12 * abs(ref), (half, w, ref)


def test_half_list_force_equals_full_after_fold():
    p, d = _setup((4, 4, 4), half=False)
    md_full = O.MD(p)
    md_full.init()
    md_full.run(10)
    ph, _ = _setup((4, 4, 4), half=True)
    md_half = O.MD(ph)
    md_half.init()
    md_half.run(10)
    xf, vf, ff, tf, idf = md_full.arrays()
    xh, vh, fh, th, idh = md_half.arrays()
    nf, nh = md_full.n_local, md_half.n_local
    of = np.argsort(idf[:nf])
    oh = np.argsort(idh[:nh])
    assert np.allclose(ff[:nf][of], fh[:nh][oh], rtol=0, atol=1e-10)
    assert np.allclose(vf[:nf][of], vh[:nh][oh], rtol=0, atol=1e-10)


# ---------------------------------------------------------- neighbours ------
def _tst_neighbor_geometry(seed=342343901):
    """tstNeighbor.hpp:250-330 geometry: 1000 atoms of which the last 200
    are ghosts, uniform in [-5.3 rc, 4.7 rc)^3, rc = 2.32 (the reference's
    Kokkos XorShift stream is replaced by numpy's)."""
    rc = 2.32
    rng = np.random.default_rng(seed)
    x = rng.uniform(-5.3 * rc, 4.7 * rc, (1000, 3))
    return x, 800, rc


def _rows(off, idx, n):
    return [idx[off[i]:off[i + 1]] for i in range(n)]


def test_neighbor_full_matches_brute_force():
    x, n_local, rc = _tst_neighbor_geometry()
    off_b, idx_b = O.neigh_brute(x, n_local, rc, half=False)
    off, idx = O.neigh_build(x, n_local, rc, half=False)
    assert len(off) == n_local + 1
    for rb, r in zip(_rows(off_b, idx_b, n_local), _rows(off, idx, n_local)):
        assert sorted(rb) == sorted(r)
    # brute force restated independently: <= rc^2, no self
    xl = x
    d2 = ((xl[:n_local, None, :] - xl[None, :, :]) ** 2).sum(-1)
    for i in range(0, n_local, 37):
        want = [j for j in range(len(x)) if j != i and d2[i, j] <= rc * rc]
        assert sorted(idx_b[off_b[i]:off_b[i + 1]]) == want


def test_neighbor_half_invariants():
    """tstNeighbor.hpp:193-243: half counts <= full counts, no pair listed
    from both sides; plus every local-local pair listed exactly once."""
    x, n_local, rc = _tst_neighbor_geometry()
    off_f, idx_f = O.neigh_build(x, n_local, rc, half=False)
    off_h, idx_h = O.neigh_build(x, n_local, rc, half=True)
    full = _rows(off_f, idx_f, n_local)
    half = _rows(off_h, idx_h, n_local)
    assert sum(map(len, half)) <= sum(map(len, full))
    pairs = set()
    for i, r in enumerate(half):
        assert len(r) <= len(full[i])
        assert set(r) <= set(full[i])
        for j in r:
            if j < n_local:
                assert i not in set(half[j])
                pairs.add((min(i, j), max(i, j)))
    want = {(min(i, j), max(i, j)) for i, r in enumerate(full) for j in r
            if j < n_local}
    assert pairs == want


def test_binning_grid_search_matches_brute_force():
    p, d = _setup((6, 6, 6))
    md = O.MD(p)
    md.init()
    x, v, f, t, ids = md.arrays()
    n = md.n_local
    gmin, rd, ng = O.bin_grid(d, [2.8, 2.8, 2.8])
    off_g, idx_g = O.neigh_build_grid(x, n, 2.8, False, gmin, rd, ng)
    off_b, idx_b = O.neigh_brute(x, n, 2.8, half=False)
    assert np.array_equal(off_g, off_b)
    for rb, r in zip(_rows(off_b, idx_b, n), _rows(off_g, idx_g, n)):
        assert sorted(rb) == sorted(r)


# ------------------------------------------------------------- binning ------
def test_bin_sort_is_stable_cell_order():
    p, d = _setup((5, 5, 5))
    x, t, ids = O.lattice(p, d)
    rng = np.random.default_rng(4)
    x = x + rng.uniform(-0.2, 0.2, x.shape)
    O.wrap(x, np.array(d.L[:]))
    perm, cell, _ = O.bin_sort(x, d, [2.8, 2.8, 2.8])
    assert sorted(perm) == list(range(len(x)))
    assert np.all(np.diff(cell) >= 0)
    # stable: within a cell the original order is kept
    for c in np.unique(cell):
        k = perm[cell == c]
        assert np.all(np.diff(k) > 0)
    # cell ids follow floor((x - min) * rd) with the cardinal (i, j, k) order
    gmin, rd, ng = O.bin_grid(d, [2.8, 2.8, 2.8])
    ijk = np.floor((x[perm] - gmin) * rd).astype(int)
    ijk = np.minimum(np.maximum(ijk, 0), np.array(ng) - 1)
    assert np.array_equal(cell, (ijk[:, 0] * ng[1] + ijk[:, 1]) * ng[2] + ijk[:, 2])


# -------------------------------------------------------------- ghosts ------
def test_ghosts_are_exactly_the_periodic_images_in_the_halo():
    p, d = _setup((5, 5, 5))
    x, t, ids = O.lattice(p, d)
    rng = np.random.default_rng(6)
    x = x + rng.uniform(-0.3, 0.3, x.shape)
    L = np.array(d.L[:])
    O.wrap(x, L)
    depth = 2.8
    xg, tg, idg, send, counts = O.ghost_build(x, t, ids, d, depth)
    n = len(x)
    ghosts = xg[n:]
    # brute force: every image shift s in {-1,0,1}^3 \ 0 whose image lies in
    # [-depth, L + depth) per dimension, where in dims with s != 0 the atom
    # is within depth of the face it is copied across
    want = set()
    for i in range(n):
        for s in itertools.product((-1, 0, 1), repeat=3):
            if s == (0, 0, 0):
                continue
            ok = True
            for k in range(3):
                if s[k] == 1 and not x[i, k] < depth:
                    ok = False
                if s[k] == -1 and not x[i, k] >= L[k] - depth:
                    ok = False
            if ok:
                img = x[i] + np.array(s) * L
                want.add((ids[i],) + tuple(np.round(img, 9)))
    got = {(idg[n + g],) + tuple(np.round(ghosts[g], 9)) for g in range(len(ghosts))}
    assert got == want
    assert sum(counts) == len(ghosts)
