"""GPU: the triangle and edge emptiness tests on the 241 known-answer
vectors of tests/golden/predicates.json ("empty_low": a probe ball e on,
just inside or just outside the smallest orthosphere of a triangle or
edge, power(e) <= rho decided with exact rationals).

Each vector becomes a 3- or 4-ball input (the simplex's balls and e),
solvent 0, alpha the smallest FP32 value at which the simplex is an alpha
simplex.  The GPU lists must equal the oracle's (whose emptiness
predicates the CPU suite pins on the same vectors), and where the output
holds no coface of the simplex -- so that only the emptiness test can keep
it -- the simplex is listed iff the vector says "not violated".  Near-ties
must reach the later precision stages on the GPU: a first-stage bound that
is too tight would decide them, and some would come out wrong.
"""
import json
import os

import numpy as np
import pytest

import pac
import oracle

pytestmark = pytest.mark.gpu


def _f32_next(a):
    return float(np.nextafter(np.float32(a), np.float32(np.inf)))


def _alpha_at(bs):
    """smallest FP32 alpha at which the simplex of bs is an alpha simplex
    (the oracle's exact filter decides), None if degenerate"""
    keep = oracle.edge_keep if len(bs) == 2 else oracle.tri_keep
    lo = float(np.float32(-1e3))
    if not keep(*bs, 1e3):
        return None
    # FP32 bisection on the ordered bit patterns
    def key(v):
        i = int(np.float32(v).view(np.int32))
        return i if i >= 0 else -(i & 0x7fffffff)

    def val(k):
        i = k if k >= 0 else (-k) | 0x80000000
        return float(np.array([i], dtype=np.uint32).view(np.float32)[0])

    a, b = key(lo), key(1e3)
    while b - a > 1:
        m = (a + b) // 2
        if keep(*bs, val(m)):
            b = m
        else:
            a = m
    return val(b)


@pytest.mark.parametrize("prec", [pac.PREC_STAGED, pac.PREC_PLAIN_FP64])
def test_emptiness_vectors_on_gpu(golden_dir, prec):
    print("case: predicates.json empty_low prec=%d" % prec, flush=True)
    d = json.load(open(os.path.join(golden_dir, "predicates.json")))
    bad, decided, escalated = [], 0, 0
    hot = 0 if prec == pac.PREC_STAGED else 1
    for v in d["empty_low"]:
        bs = [[float.fromhex(x) for x in b] for b in v["balls"]]
        e = [float.fromhex(x) for x in v["e"]]
        al = _alpha_at(bs)
        assert al is not None
        balls = np.array(bs + [e], dtype=np.float32)
        k = len(bs) - 1  # 1 edge, 2 triangle
        fam = 5 if k == 1 else 4
        r = pac.compute(balls, 0.0, al, precision=prec)
        g = r.lists()
        o, _, _ = oracle.run(balls, 0.0, al)
        for dd in (1, 2, 3):
            if not np.array_equal(g[dd], o[dd]):
                bad.append(("oracle", k, v["violated"], dd))
        st = r.stats()
        escalated += sum(int(st.stage_evals[fam][s]) for s in range(hot + 1, 4))
        simplex = tuple(range(len(bs)))
        cofaces = [row for dd in range(k + 1, 4) for row in g[dd].tolist()
                   if set(simplex) <= set(row)]
        if not cofaces:
            decided += 1
            listed = any(tuple(sorted(row)) == simplex for row in g[k].tolist())
            if listed != (not v["violated"]):
                bad.append(("emptiness", k, v["violated"], listed))
    assert not bad, bad[:5]
    # most vectors are settled by the emptiness test alone
    assert decided > len(d["empty_low"]) // 2
    # near-ties go past the first stage
    assert escalated > 0
