"""CPU: simulation of simplicity for the weighted insphere determinant.

Pins: the term table derived from the Edelsbrunner-Muecke definition (49
terms, the same minors and flip set as the reference's hand-written table,
predicates.h:4719-5447) in three independent places -- the Python
derivation, the oracle's C++ enumeration, and the table compiled into the
device code; the sign conventions insphere_w == det[x y z l 1] and
sign(insphere*orient) == sign(rho - power); and, end to end, that SoS
resolves a fully co-spherical input (a cubic lattice with equal radii) into
a valid triangulation."""
import itertools
import os
import sys
from fractions import Fraction as F

import numpy as np
import pytest

import oracle
import pac

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "parallelac-xmp_amd", "csrc"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import gen_sos_table  # noqa: E402


def test_tables_agree():
    py = np.array(gen_sos_table.terms(), dtype=np.int32)
    orc = oracle.sos_table()
    import ctypes
    L = pac.lib()
    L.pac_internal_sos_table.argtypes = [ctypes.c_void_p, ctypes.c_int]
    buf = np.zeros((64, 3), dtype=np.int32)
    n = L.pac_internal_sos_table(buf.ctypes.data, 64)
    dev = buf[:n]
    assert len(py) == 49
    assert np.array_equal(py, orc)
    assert np.array_equal(py, dev)
    # first term: orient3d of points 1..4 (non-lifted), negative coefficient
    assert tuple(py[0]) == (0x1E, 0x17, -1)
    # every term keeps the column of ones (the perturbation never touches
    # it) and is square: the device evaluates each m x m minor as the
    # (m-1) x (m-1) determinant of row differences
    for rm, cm, _ in py:
        assert cm & 0x10
        assert bin(rm).count("1") == bin(cm).count("1")
    # flip set of the reference's table (predicates.h:5420-5447)
    flips = {i + 1 for i, t in enumerate(py) if t[2] < 0}
    assert flips == {1, 3, 7, 9, 11, 14, 15, 16, 18, 20, 22, 23, 26, 30, 31, 32,
                     37, 39, 41, 44, 46}


def _rand_ball(rng):
    return [float(np.float32(round(rng.uniform(-3, 3), 3))) for _ in range(3)] + \
           [float(np.float32(rng.choice([1.7, 1.55, 1.52]) + 1.4))]


def test_sos_equals_sign_when_nondegenerate():
    rng = np.random.default_rng(3)
    for _ in range(200):
        b = [_rand_ball(rng) for _ in range(5)]
        ids = list(rng.permutation(1000)[:5])
        d = oracle.det5(b)
        if d != 0:
            assert oracle.sos_det5(b, ids) == d


def test_sos_never_zero_and_alternating_on_lattice():
    pts = [[x, y, z, 1.7] for x in (0.0, 2.0) for y in (0.0, 2.0) for z in (0.0, 2.0)]
    rng = np.random.default_rng(5)
    for _ in range(300):
        idx = list(rng.choice(8, 5, replace=False))
        b = [pts[i] for i in idx]
        ids = list(rng.permutation(100)[:5])
        s = oracle.sos_det5(b, ids)
        assert s in (-1, 1)
        # transposition of two rows (with their ids) flips the sign
        i, j = rng.choice(5, 2, replace=False)
        b2, ids2 = list(b), list(ids)
        b2[i], b2[j] = b2[j], b2[i]
        ids2[i], ids2[j] = ids2[j], ids2[i]
        assert oracle.sos_det5(b2, ids2) == -s


def test_orient_and_insphere_conventions():
    rng = np.random.default_rng(11)
    for _ in range(100):
        b = [_rand_ball(rng) for _ in range(5)]
        o = oracle.orient3d(*b[:4])
        if o == 0:
            continue
        d = oracle.det5(b)
        P = [[F(v) for v in x[:3]] for x in b]
        W = [F(x[3]) ** 2 for x in b]
        a = P[0]
        import make_golden  # noqa: F401
        from make_golden import ortho, power
        c, rho = ortho(b[:4])
        pe = power(c, b[4])
        sgn = (rho > pe) - (rho < pe)
        assert d * o == sgn


def _vol6(p):
    a, b, c, d = [np.array(x[:3], dtype=object) for x in p]
    u, v, w = b - a, c - a, d - a
    return (u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0])
            + u[2] * (v[0] * w[1] - v[1] * w[0]))


def test_lattice_triangulation_is_valid():
    """4x4x4 lattice, spacing 2, equal radii, alpha large: every tet is an
    alpha tet and all 8-point cubes are co-spherical.  The kept tets must
    tile the cube exactly: |volumes| sum to the cube volume, every interior
    triangle has exactly two tets, every boundary triangle one."""
    m = 4
    g = [[2.0 * x, 2.0 * y, 2.0 * z, 1.7] for z in range(m) for y in range(m) for x in range(m)]
    b = np.array(g, dtype=np.float32)
    out, st, _ = oracle.run(b, 0.0, 200.0)
    tets = out[3]
    assert st["sos"] > 0
    vol = sum(abs(_vol6([g[i] for i in t])) for t in tets)
    side = 2.0 * (m - 1)
    assert vol == 6 * side ** 3
    assert all(_vol6([g[i] for i in t]) != 0 for t in tets)
    from collections import Counter
    fc = Counter(tuple(sorted(f)) for t in tets for f in itertools.combinations(t, 3))

    def on_boundary(f):
        P = np.array([g[i][:3] for i in f])
        return any((P[:, a] == 0).all() or (P[:, a] == side).all() for a in range(3))
    for f, k in fc.items():
        assert k == (1 if on_boundary(f) else 2), (f, k)
