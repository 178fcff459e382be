"""Writes tests/golden/predicates.json: known-answer vectors for the exact
predicates, computed independently with Python Fractions (exact rational
arithmetic on the FP32 inputs).

The orthocentre of each simplex is obtained by solving the linear system
of its radical planes by Gaussian elimination (not by the closed forms the
oracle and the kernels use), so a dropped term, wrong sign or transposed
operand in those closed forms fails these vectors.

Vectors deliberately sit on and next to the decision boundaries:
  - filters: alpha = rho rounded to FP32 and its two FP32 neighbours, so
    rho <= alpha is exercised at exact ties (reference tie rule '<=' with
    ATHRES = 0, AlphaComplexLargeData.cu:49, :187, :347, :589);
  - triangle / edge emptiness: a ball placed on the orthosphere (power
    exactly rho after FP32 rounding in some cases) and nudged in/out
    (reference '<=' with THRES = 0, :838, :971).

Run:  python3 tests/golden/make_golden.py
"""
import json
import os
import random
from fractions import Fraction as F

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def f32(x):
    return float(np.float32(x))


def fx(v):
    return F(float(v))


def solve(A, b):
    n = len(A)
    M = [list(A[i]) + [b[i]] for i in range(n)]
    for c in range(n):
        p = next((r for r in range(c, n) if M[r][c] != 0), None)
        if p is None:
            return None
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [M[r][k] - f * M[c][k] for k in range(n + 1)]
    return [M[i][n] / M[i][i] for i in range(n)]


def ortho(balls):
    """Smallest orthosphere (centre, rho) of 2-4 balls (x,y,z,r)."""
    P = [[fx(b[0]), fx(b[1]), fx(b[2])] for b in balls]
    W = [fx(b[3]) ** 2 for b in balls]
    a = P[0]
    rows, rhs = [], []
    for q, w in zip(P[1:], W[1:]):
        rows.append([2 * (q[i] - a[i]) for i in range(3)])
        rhs.append(sum((q[i] - a[i]) ** 2 for i in range(3)) + W[0] - w)
    k = len(balls)
    if k < 4:
        # restrict x to the affine span: x = sum t_m (q_m - a)
        U = [[q[i] - a[i] for i in range(3)] for q in P[1:]]
        A = [[sum(2 * U[r][i] * U[m][i] for i in range(3)) for m in range(k - 1)]
             for r in range(k - 1)]
        t = solve(A, rhs)
        if t is None:
            return None
        x = [sum(t[m] * U[m][i] for m in range(k - 1)) for i in range(3)]
    else:
        x = solve(rows, rhs)
        if x is None:
            return None
    c = [a[i] + x[i] for i in range(3)]
    rho = sum(xx * xx for xx in x) - W[0]
    return c, rho


def power(c, b):
    return sum((fx(b[i]) - c[i]) ** 2 for i in range(3)) - fx(b[3]) ** 2


def rand_ball(rng, cx=0.0, spread=2.5):
    return [f32(round(cx + rng.uniform(-spread, spread), 3)) for _ in range(3)] + [
        f32(np.float32(rng.choice([1.70, 1.55, 1.52, 1.80])) + np.float32(1.4))]


def main():
    rng = random.Random(20261202)
    out = {"filters": [], "empty_low": [], "note": "balls are (x,y,z,r) FP32 "
           "with r already including the solvent; values stored as float.hex"}
    for kind in (2, 3, 4):
        made = 0
        while made < 40:
            balls = [rand_ball(rng) for _ in range(kind)]
            o = ortho(balls)
            if o is None:
                continue
            c, rho = o
            a0 = f32(float(rho))
            for alpha in (float(np.nextafter(np.float32(a0), np.float32(-np.inf))), a0,
                          float(np.nextafter(np.float32(a0), np.float32(np.inf)))):
                keep = int(F(alpha) >= rho)
                out["filters"].append({"balls": [[float(v).hex() for v in b] for b in balls],
                                       "alpha": float(alpha).hex(), "keep": keep,
                                       "rho": float(rho)})
            made += 1
    # degenerate filters: coincident / collinear / coplanar centres -> reject
    b0 = [0.0, 0.0, 0.0, 2.9]
    out["filters"].append({"balls": [[v.hex() for v in b0], [v.hex() for v in b0]],
                           "alpha": (10.0).hex(), "keep": 0, "rho": None})
    col = [[0.0, 0.0, 0.0, 3.0], [1.0, 1.0, 1.0, 3.0], [2.5, 2.5, 2.5, 3.0]]
    out["filters"].append({"balls": [[float(v).hex() for v in b] for b in col],
                           "alpha": (10.0).hex(), "keep": 0, "rho": None})
    cop = [[0.0, 0.0, 0.0, 3.0], [2.0, 0.0, 0.0, 3.0], [0.0, 2.0, 0.0, 3.0],
           [2.0, 2.0, 0.0, 3.0]]
    out["filters"].append({"balls": [[float(v).hex() for v in b] for b in cop],
                           "alpha": (10.0).hex(), "keep": 0, "rho": None})
    # emptiness of triangles / edges: probe ball near the orthosphere
    for kind in (2, 3):
        made = 0
        while made < 40:
            balls = [rand_ball(rng) for _ in range(kind)]
            o = ortho(balls)
            if o is None:
                continue
            c, rho = o
            # direction and radius for the probe: power(c, e) = rho
            re_ = f32(np.float32(1.55) + np.float32(1.4))
            d2 = rho + fx(re_) ** 2
            if d2 <= 0:
                continue
            dist = float(d2) ** 0.5
            u = [rng.gauss(0, 1) for _ in range(3)]
            nu = sum(x * x for x in u) ** 0.5
            base = [float(c[i]) + dist * u[i] / nu for i in range(3)]
            for scale in (0.999, 1.0, 1.001):
                e = [f32(float(c[i]) + (base[i] - float(c[i])) * scale) for i in range(3)] + [re_]
                viol = int(power(c, e) <= rho)
                out["empty_low"].append({"balls": [[float(v).hex() for v in b] for b in balls],
                                         "e": [float(v).hex() for v in e], "violated": viol})
            made += 1
    # exact tie: lattice square + probe on the circle (power == rho exactly)
    sq = [[0.0, 0.0, 0.0, 1.7], [2.0, 0.0, 0.0, 1.7], [0.0, 2.0, 0.0, 1.7]]
    e = [2.0, 2.0, 0.0, 1.7]
    o = ortho(sq)
    out["empty_low"].append({"balls": [[float(v).hex() for v in b] for b in sq],
                             "e": [float(v).hex() for v in e],
                             "violated": int(power(o[0], e) <= o[1])})
    with open(os.path.join(HERE, "predicates.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("filters", len(out["filters"]), "empty_low", len(out["empty_low"]))


if __name__ == "__main__":
    main()
