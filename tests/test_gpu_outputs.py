"""GPU: alpha values (rho per output simplex, FP64) against exact rationals,
and the canonical row order against the reference order."""
import os
import sys
from fractions import Fraction as F

import numpy as np
import pytest

import pac

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import ortho  # noqa: E402

pytestmark = pytest.mark.gpu

# FP64 evaluation of the closed forms from FP32 inputs; error relative to the
# magnitude of the terms (|rho| + w of the first vertex).
RHO_REL_TOL = {1: 1e-12, 2: 1e-10, 3: 1e-9}


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_alpha_values_vs_exact(dim):
    print("case: alpha values dim", dim, flush=True)
    b = pac.synth_globule(1000, 16.84, 1.3, seed=1)
    r = pac.compute(b, 1.4, 0.0)
    rows = r.lists()[dim]
    rho = r.alpha_values(dim)
    assert rho.shape == (len(rows),)
    balls = [[float(x), float(y), float(z), float(np.float32(np.float32(w) + np.float32(1.4)))]
             for x, y, z, w in b]
    worst = 0.0
    idx = np.linspace(0, len(rows) - 1, 300).astype(int)
    for i in idx:
        c, ex = ortho([balls[v] for v in rows[i]])
        scale = abs(float(ex)) + balls[rows[i][0]][3] ** 2
        worst = max(worst, abs(rho[i] - float(ex)) / scale)
        # every output simplex is an alpha simplex: rho <= alpha = 0
        assert ex <= 0
    print("dim %d max relative error %.3e" % (dim, worst))
    assert worst <= RHO_REL_TOL[dim]


def test_canonical_order():
    print("case: canonical order", flush=True)
    b = pac.synth_globule(2000, None, 1.3, seed=4)
    ref = pac.compute(b, 1.4, 0.0).lists()
    can = pac.compute(b, 1.4, 0.0, order=pac.ORDER_CANONICAL).lists()
    for d in (1, 2, 3):
        exp = np.sort(ref[d], axis=1)
        exp = exp[np.lexsort(exp.T[::-1])]
        assert np.array_equal(can[d], exp)
