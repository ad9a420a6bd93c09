"""NNP model descriptions shared by the NNP tests.

- random_model(): a small seeded two-element Behler-Parrinello model.
- read_n2p2(): an independent reader of n2p2 model files (input.nn,
  scaling.data, weights.NNN.data), used to cross-check the engine's loader
  and to run the reference's fixture (input/nnp, copied to tests/golden/nnp_ni).
- cluster(): a seeded open cluster of atoms with a minimum separation.

A model is a dict: nelem, hidden, act, smin, smax, scale, sf[e] = list of
(type, e1, e2, eta, rs, lambda, zeta, rc), gmin[e], gmax[e], W[e][l] (out x in),
b[e][l], atom_energy[e].
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "nnp_ni")


def random_model(seed=3, rc=4.0, hidden=(6, 5)):
    rng = np.random.default_rng(seed)
    sf = []
    for e in range(2):
        lst = []
        for eta in (0.05, 0.2, 0.6):
            for ne in (0, 1):
                lst.append((2, ne, -1, eta, 0.5 * e, 0.0, 0.0, rc))
        for eta in (0.01, 0.1):
            for zeta in (1.0, 2.0, 4.0):
                for lam in (-1.0, 1.0):
                    lst.append((3, -1, -1, eta, 0.0, lam, zeta, rc))
        lst.append((3, 0, 1, 0.05, 0.0, 1.0, 1.0, rc))
        sf.append(lst)
    m = dict(nelem=2, hidden=list(hidden), act="ttl", smin=0.0, smax=1.0,
             scale=True, sf=sf, gmin=[], gmax=[], W=[], b=[],
             atom_energy=[-0.5, -0.25])
    for e in range(2):
        n = len(sf[e])
        gmin = rng.uniform(0.0, 0.2, n)
        m["gmin"].append(gmin)
        m["gmax"].append(gmin + rng.uniform(0.5, 2.0, n))
        sizes = [n] + list(hidden) + [1]
        Ws, bs = [], []
        for l in range(len(sizes) - 1):
            Ws.append(rng.normal(0, 1, (sizes[l + 1], sizes[l])) /
                      np.sqrt(sizes[l]))
            bs.append(rng.normal(0, 0.1, sizes[l + 1]))
        m["W"].append(Ws)
        m["b"].append(bs)
    return m


def cluster(n=24, box=5.0, dmin=0.9, seed=5, ntypes=2):
    rng = np.random.default_rng(seed)
    x = []
    while len(x) < n:
        p = rng.uniform(0, box, 3)
        if all(np.linalg.norm(p - q) >= dmin for q in x):
            x.append(p)
    x = np.array(x)
    t = rng.integers(0, ntypes, n).astype(np.int32)
    return x, t


def to_oracle(O, m):
    M = O.NNP(m["nelem"], m["hidden"], m["act"], m["scale"], m["smin"],
              m["smax"])
    for e in range(m["nelem"]):
        for k, s in enumerate(m["sf"][e]):
            M.add_sf(e, *s, gmin=float(m["gmin"][e][k]),
                     gmax=float(m["gmax"][e][k]))
        for l in range(len(m["W"][e])):
            M.set_layer(e, l, m["W"][e][l], m["b"][e][l])
        M.set_atom_energy(e, m["atom_energy"][e])
    return M


def _sf_sort_key(s):
    # radial before angular; then cutoff, eta, rs, zeta, lambda, elements
    typ, e1, e2, eta, rs, lam, zeta, rc = s
    return (typ, rc, eta, rs, zeta, lam, e1, e2)


def read_n2p2(path=FIXTURE, elements=("Ni",), atomic_numbers=(28,)):
    """Parse an n2p2 model directory (the keywords the fixture uses)."""
    sf = [[] for _ in elements]
    hidden, act = [], ""
    smin, smax, scale = 0.0, 1.0, False
    atom_energy = [0.0] * len(elements)
    with open(os.path.join(path, "input.nn")) as fh:
        for line in fh:
            w = line.split("#")[0].split()
            if not w:
                continue
            if w[0] == "global_nodes_short":
                hidden = [int(v) for v in w[1:]]
            elif w[0] == "global_activation_short":
                act = "".join(w[1:])
            elif w[0] == "scale_symmetry_functions":
                scale = True
            elif w[0] == "scale_min_short":
                smin = float(w[1])
            elif w[0] == "scale_max_short":
                smax = float(w[1])
            elif w[0] == "atom_energy":
                atom_energy[elements.index(w[1])] = float(w[2])
            elif w[0] == "symfunction_short":
                e = elements.index(w[1])
                typ = int(w[2])
                if typ == 2:
                    e1 = elements.index(w[3])
                    eta, rs, rc = float(w[4]), float(w[5]), float(w[6])
                    sf[e].append((2, e1, -1, eta, rs, 0.0, 0.0, rc))
                elif typ == 3:
                    e1, e2 = elements.index(w[3]), elements.index(w[4])
                    eta, lam, zeta, rc = (float(w[5]), float(w[6]),
                                          float(w[7]), float(w[8]))
                    rs = float(w[9]) if len(w) > 9 else 0.0
                    sf[e].append((3, e1, e2, eta, rs, lam, zeta, rc))
    sf = [sorted(s, key=_sf_sort_key) for s in sf]
    gmin = [np.zeros(len(s)) for s in sf]
    gmax = [np.zeros(len(s)) for s in sf]
    with open(os.path.join(path, "scaling.data")) as fh:
        for line in fh:
            w = line.split()
            if not w or w[0].startswith("#"):
                continue
            e, k = int(w[0]) - 1, int(w[1]) - 1
            gmin[e][k], gmax[e][k] = float(w[2]), float(w[3])
    W, b = [], []
    nparams = 0
    for e, z in enumerate(atomic_numbers):
        sizes = [len(sf[e])] + hidden + [1]
        We = [np.zeros((sizes[l + 1], sizes[l])) for l in range(len(sizes) - 1)]
        be = [np.zeros(sizes[l + 1]) for l in range(len(sizes) - 1)]
        with open(os.path.join(path, "weights.%03d.data" % z)) as fh:
            for line in fh:
                w = line.split()
                if not w or w[0].startswith("#"):
                    continue
                nparams += 1
                v = float(w[0])
                if w[1] == "a":
                    lf, nf, lt, nt = (int(w[3]), int(w[4]), int(w[5]),
                                      int(w[6]))
                    We[lf][nt - 1, nf - 1] = v
                else:
                    l, nd = int(w[3]), int(w[4])
                    be[l - 1][nd - 1] = v
        W.append(We)
        b.append(be)
    return dict(nelem=len(elements), hidden=hidden, act=act, smin=smin,
                smax=smax, scale=scale, sf=sf, gmin=gmin, gmax=gmax, W=W,
                b=b, atom_energy=atom_energy, nparams=nparams)


def sf_brute(m, x, t, rc_list=None):
    """Independent O(N^3) evaluation of every atom's symmetry functions of an
    open cluster (no periodic images), straight from the definitions."""
    n = len(x)
    out = []
    for i in range(n):
        lst = m["sf"][t[i]]
        g = np.zeros(len(lst))
        for k, (typ, e1, e2, eta, rs, lam, zeta, rc) in enumerate(lst):
            fc = lambda r: 0.5 * (np.cos(np.pi * r / rc) + 1.0) if r < rc else 0.0
            if typ == 2:
                for j in range(n):
                    if j == i or (e1 >= 0 and t[j] != e1):
                        continue
                    r = np.linalg.norm(x[j] - x[i])
                    g[k] += np.exp(-eta * (r - rs) ** 2) * fc(r)
            else:
                for j in range(n):
                    for q in range(j + 1, n):
                        if i in (j, q):
                            continue
                        if e1 >= 0 and {t[j], t[q]} != {e1, e2} and not (
                                e1 == e2 and t[j] == e1 and t[q] == e1):
                            continue
                        a, bb = x[j] - x[i], x[q] - x[i]
                        r1, r2 = np.linalg.norm(a), np.linalg.norm(bb)
                        r3 = np.linalg.norm(bb - a)
                        c = a.dot(bb) / (r1 * r2)
                        g[k] += (2.0 ** (1 - zeta) * (1 + lam * c) ** zeta *
                                 np.exp(-eta * ((r1 - rs) ** 2 + (r2 - rs) ** 2 +
                                                (r3 - rs) ** 2)) *
                                 fc(r1) * fc(r2) * fc(r3))
        out.append(g)
    return out
