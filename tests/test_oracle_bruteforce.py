"""CPU: the oracle pipeline against a brute-force statement of the weighted
alpha complex on tiny inputs, in exact rational arithmetic:
  alpha simplex   = every 2/3/4-subset whose smallest orthosphere has
                    rho <= alpha (non-degenerate);
  tet kept        = no other ball with power < rho (generic inputs: no ties);
  triangle kept   = face of a kept tet, or no other ball with power <= rho;
  edge kept       = edge of a kept triangle, or no other ball with power <= rho.
This is the semantics of AlphaComplexParallel.cu:1449-1788 with exact
decisions; it needs no grid, so it also pins the grid-independence of the
exact cell policy."""
import itertools
import os
import sys
from fractions import Fraction as F

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import ortho, power  # noqa: E402

import oracle


def brute(b, s, alpha):
    n = len(b)
    balls = [[float(x), float(y), float(z), float(np.float32(np.float32(r) + np.float32(s)))]
             for x, y, z, r in b]
    A = F(float(np.float32(alpha)))
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
