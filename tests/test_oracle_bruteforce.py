"""CPU: the oracle pipeline against a brute-force statement of the weighted
alpha complex on tiny inputs, in exact rational arithmetic:
  alpha simplex   = every 2/3/4-subset whose smallest orthosphere has
                    rho <= alpha (non-degenerate);
  tet kept        = no other ball with power < rho (generic inputs: no ties);
  triangle kept   = face of a kept tet, or no other ball with power <= rho;
  edge kept       = edge of a kept triangle, or no other ball with power <= rho.
This is the semantics of AlphaComplexParallel.cu:1449-1788 with exact
decisions; it needs no grid, so it also pins the grid-independence of the
exact cell policy.  It runs on FP32 balls and on continuous FP64 balls
(the oracle's FP64-input path, whose coordinates near 0 carry bits down to
2^-69), and an exact emulation of the reference's own grid and 27-cell
stencils pins the oracle's reference_raw cell policy."""
import itertools
import os
import sys
from fractions import Fraction as F

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import ortho, power  # noqa: E402

import oracle


def grown(b, s):
    """Balls with the solvent added in the input's own float type:
    fl(r_raw + s) (AlphaComplexParallel.cu:1086 for FP32)."""
    if np.asarray(b).dtype == np.float64:
        return [[float(x), float(y), float(z), float(np.float64(r) + np.float64(s))]
                for x, y, z, r in b]
    return [[float(x), float(y), float(z), float(np.float32(np.float32(r) + np.float32(s)))]
            for x, y, z, r in b]


def brute(b, s, alpha):
    n = len(b)
    balls = grown(b, s)
    A = F(float(alpha if np.asarray(b).dtype == np.float64 else np.float32(alpha)))
    al = {2: {}, 3: {}, 4: {}}
    for k in (2, 3, 4):
        for sub in itertools.combinations(range(n), k):
            o = ortho([balls[i] for i in sub])
            if o is None:
                continue
            c, rho = o
            if rho <= A:
                al[k][sub] = (c, rho)
    tets = set()
    for sub, (c, rho) in al[4].items():
        if all(power(c, balls[e]) >= rho for e in range(n) if e not in sub) and \
           all(power(c, balls[e]) > rho for e in range(n) if e not in sub):
            tets.add(sub)
    tris = set()
    faces = {f for t in tets for f in itertools.combinations(t, 3)}
    for sub, (c, rho) in al[3].items():
        if sub in faces or all(power(c, balls[e]) > rho for e in range(n) if e not in sub):
            tris.add(sub)
    edges = set()
    tedges = {e for t in tris for e in itertools.combinations(t, 2)}
    for sub, (c, rho) in al[2].items():
        if sub in tedges or all(power(c, balls[e]) > rho for e in range(n) if e not in sub):
            edges.add(sub)
    return edges, tris, tets


@pytest.mark.parametrize("seed", range(6))
def test_bruteforce_small(seed):
    rng = np.random.default_rng(seed)
    n = 14
    pts = np.round(rng.uniform(-4.0, 4.0, size=(n, 3)), 3)
    rad = rng.choice([1.70, 1.55, 1.52, 1.80], size=(n, 1))
    b = np.concatenate([pts, rad], 1).astype(np.float32)
    s = [0.0, 1.4][seed % 2]
    alpha = [0.0, 1.0, -0.5][seed % 3]
    E, T, K = brute(b, s, alpha)
    for policy in (0,):
        out, st, _ = oracle.run(b, s, alpha, cell_policy=policy)
        got = [set(tuple(sorted(r)) for r in out[d].tolist()) for d in (1, 2, 3)]
        assert got[0] == E
        assert got[1] == T
        assert got[2] == K


def lsb_exp(v):
    """Exponent of the lowest set bit of a nonzero double."""
    m, e = np.frexp(v)
    m = int(np.ldexp(abs(m), 53))
    return int(e) - 53 + ((m & -m).bit_length() - 1)


def f64_balls(seed, n=14):
    """Continuous FP64 centres (uniform + N(0, 0.01) jitter, no rounding);
    a few coordinates are put within 1e-5..1e-3 of 0, so their lowest set
    bits sit at 2^-62..2^-69 and the oracle's exact embedding needs its
    largest scale."""
    rng = np.random.default_rng(1000 + seed)
    pts = rng.uniform(-2.5, 2.5, size=(n, 3)) + rng.normal(0.0, 0.01, size=(n, 3))
    for q in range(3):
        pts[q, q] = rng.choice([-1.0, 1.0]) * 10.0 ** rng.uniform(-5.0, -3.0)
    rad = rng.choice([1.70, 1.55, 1.52, 1.80], size=(n, 1))
    return np.concatenate([pts, rad], 1).astype(np.float64)


@pytest.mark.parametrize("seed", range(8))
def test_bruteforce_f64(seed):
    """The oracle's FP64-input path (pac_oracle_run_f64: FP64 balls, exact
    scaled integers) against the exact rational brute force.  Fraction(float)
    is exact for doubles, so this pins that path independently."""
    b = f64_balls(seed)
    assert min(lsb_exp(v) for v in b[:, :3].ravel() if v != 0) <= -60
    s = (0.0, 1.4)[seed % 2]
    alpha = (0.0, 0.8)[(seed // 2) % 2]
    E, T, K = brute(b, s, alpha)
    assert len(K) > 0
    out, st, _ = oracle.run(b, s, alpha, cell_policy=0)
    assert st["scale"] >= 60
    got = [set(tuple(sorted(r)) for r in out[d].tolist()) for d in (1, 2, 3)]
    assert got[0] == E
    assert got[1] == T
    assert got[2] == K


# ---------------------------------------------------------------------------
# The reference's own grid and stencils (cell_policy = reference_raw), with
# exact decisions: an independent emulation of AlphaComplexParallel.cu.
# ---------------------------------------------------------------------------

def ref_grid(b, alpha):
    """FP32 grid of AlphaComplexParallel.cu:1107-1124 (step = 2 sqrtf(maxR^2 +
    alpha) from the raw radii, extents (int)((max-min)/step)+1) and the FP32
    cell key of :107-120.  FP64 balls are keyed by their FP32 roundings."""
    f = np.float32
    a = np.asarray(b)
    xyz = a[:, :3].astype(np.float32)
    mn = xyz.min(0)
    mx = xyz.max(0)
    maxr = f(a[:, 3].max())
    m2a = f(f(maxr * maxr) + f(alpha))
    step = f(f(2.0) * np.sqrt(m2a))
    ext = [int(f(f(mx[q] - mn[q]) / step)) + 1 for q in range(3)]
    cells = np.zeros((len(a), 3), dtype=np.int64)
    for i in range(len(a)):
        for q in range(3):
            cells[i, q] = int(f(f(xyz[i, q] - mn[q]) / step))
    key = cells[:, 0] + ext[0] * (cells[:, 1] + ext[1] * cells[:, 2])
    return mn, step, ext, cells, key


def brute_raw(b, s, alpha):
    """Lists (reference row order, original ids) of the reference pipeline on
    its own grid with exact decisions:
      - sorted ids: stable sort by cell key (:1154-1160);
      - candidate edges (i < j sorted) only between cells adjacent in every
        axis (27-cell stencil, :122-201); triangles and tets are cliques of
        alpha edges / alpha triangles (:232-308, :418-543);
      - emptiness scans only the 27 cells around the cell of the exact
        orthocentre, clamped to the grid (:611-749, :807-1015);
      - closure as :1449-1788."""
    balls0 = grown(b, s)
    A = F(float(alpha if np.asarray(b).dtype == np.float64 else np.float32(alpha)))
    mn, step, ext, cells, key = ref_grid(b, alpha)
    perm = np.argsort(key, kind="stable")
    n = len(b)
    B = [balls0[perm[q]] for q in range(n)]
    C = [cells[perm[q]] for q in range(n)]
    fmn = [F(float(v)) for v in mn]
    fstep = F(float(step))

    def near(i, j):
        return all(abs(int(C[i][q]) - int(C[j][q])) <= 1 for q in range(3))

    def scan(c):
        lo, hi = [], []
        for q in range(3):
            cq = (c[q] - fmn[q]) / fstep
            fl = cq.numerator // cq.denominator
            lo.append(max(0, fl - 1))
            hi.append(min(ext[q] - 1, fl + 1))
        return [e for e in range(n)
                if all(lo[q] <= C[e][q] <= hi[q] for q in range(3))]

    al = {2: {}, 3: {}, 4: {}}
    for i, j in itertools.combinations(range(n), 2):
        if near(i, j):
            o = ortho([B[i], B[j]])
            if o is not None and o[1] <= A:
                al[2][(i, j)] = o
    for i, j, k in itertools.combinations(range(n), 3):
        if (i, j) in al[2] and (i, k) in al[2] and (j, k) in al[2]:
            o = ortho([B[i], B[j], B[k]])
            if o is not None and o[1] <= A:
                al[3][(i, j, k)] = o
    for t in itertools.combinations(range(n), 4):
        if all(f in al[3] for f in itertools.combinations(t, 3)):
            o = ortho([B[v] for v in t])
            if o is not None and o[1] <= A:
                al[4][t] = o
    tets = set()
    for t, (c, rho) in al[4].items():
        pw = [power(c, B[e]) for e in scan(c) if e not in t]
        assert all(p != rho for p in pw), "exact tie: needs SoS"
        if all(p > rho for p in pw):
            tets.add(t)
    faces = {f for t in tets for f in itertools.combinations(t, 3)}
    tris = set(faces)
    for t, (c, rho) in al[3].items():
        if t not in faces and all(power(c, B[e]) > rho for e in scan(c) if e not in t):
            tris.add(t)
    tedges = {e for t in tris for e in itertools.combinations(t, 2)}
    edges = set(tedges)
    for t, (c, rho) in al[2].items():
        if t not in tedges and all(power(c, B[e]) > rho for e in scan(c) if e not in t):
            edges.add(t)
    return [np.array([[perm[v] for v in row] for row in sorted(S)], dtype=np.int64).reshape(-1, d + 1)
            for d, S in ((1, edges), (2, tris), (3, tets))]


def globule(seed, n, f64=False):
    """Poisson-disk (min 1.3 A) centres in a sphere at 0.05 balls/A^3, the
    four-radius mix, coordinates rounded to 0.001 A (FP32) or jittered FP64."""
    rng = np.random.default_rng(2000 + seed)
    R = (3 * n / (4 * np.pi * 0.05)) ** (1 / 3)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-R, R, 3)
        if p @ p <= R * R and all(((p - q) @ (p - q)) >= 1.69 for q in pts):
            pts.append(p)
    pts = np.array(pts)
    rad = rng.choice([1.70, 1.55, 1.52, 1.80], size=(n, 1), p=[0.63, 0.17, 0.19, 0.01])
    if f64:
        return np.concatenate([pts + rng.normal(0, 0.01, pts.shape), rad], 1)
    return np.concatenate([np.round(pts, 3), rad], 1).astype(np.float32)


RAW_CASES = [(0, 1.4, 0.0, False), (1, 1.4, 0.0, False), (2, 0.0, 1.0, False),
             (3, 1.4, 0.5, False), (4, 2.0, 0.0, False), (5, 1.4, 0.0, True)]


@pytest.mark.parametrize("seed,s,alpha,f64", RAW_CASES)
def test_reference_raw_emulation(seed, s, alpha, f64):
    """oracle.run(cell_policy=1) against the exact emulation of the
    reference's grid and 27-cell stencils, row for row."""
    b = globule(seed, 40, f64)
    want = brute_raw(b, s, alpha)
    out, st, _ = oracle.run(b, s, alpha, cell_policy=1)
    assert st["k_enum"] == 1 and st["k_ac2"] == 1
    for d in (1, 2, 3):
        np.testing.assert_array_equal(out[d], want[d - 1])


def test_reference_raw_differs_from_exact():
    """Non-vacuous: at solvent 1.4 the reference's raw-radius grid misses
    simplices of the exact complex (its stencil is too small), and the
    emulation reproduces that.  The exact complex here is the oracle's
    exact cell policy, which the brute force above pins."""
    diff = 0
    for seed, s, alpha, f64 in RAW_CASES[:2]:
        b = globule(seed, 40, f64)
        raw = brute_raw(b, s, alpha)
        ex, _, _ = oracle.run(b, s, alpha, cell_policy=0)
        for d in (1, 2, 3):
            a = set(map(tuple, np.sort(raw[d - 1], 1).tolist()))
            diff += len(a ^ set(map(tuple, np.sort(ex[d], 1).tolist())))
    assert diff > 0
