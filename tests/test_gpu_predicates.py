"""GPU: the alpha filters on the known-answer vectors of
tests/golden/predicates.json (exact rational arithmetic; alpha at rho
rounded to FP32 and at its two FP32 neighbours, so every vector sits on
the decision threshold).  Each vector's 2, 3 or 4 balls are the whole
input: the simplex is in the output iff it is an alpha simplex (no other
ball can violate it).  A first-stage bound that is too tight decides such
a vector in FP32 or FP64 and gets it wrong; a correct ladder escalates it.
"""
import json
import os

import numpy as np
import pytest

import pac

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [pac.PREC_STAGED, pac.PREC_PLAIN_FP64])
def test_filter_vectors_on_gpu(golden_dir, prec):
    print("case: predicates.json filters prec=%d" % prec, flush=True)
    d = json.load(open(os.path.join(golden_dir, "predicates.json")))
    bad, escalated = [], 0
    hot = 0 if prec == pac.PREC_STAGED else 1
    for v in d["filters"]:
        b = np.array([[float.fromhex(x) for x in ball] for ball in v["balls"]],
                     dtype=np.float32)
        al = float.fromhex(v["alpha"])
        k = len(b) - 1
        try:
            r = pac.compute(b, 0.0, al, precision=prec)
        except pac.PacError as e:
            # grid step 2 sqrt(maxR^2 + alpha) not positive: alpha is below
            # every ball's -r^2, so nothing can be kept
            assert e.code == pac.PAC_ERR_INVALID
            got = 0
        else:
            got = int(len(r.lists()[k]) == 1)
            st = r.stats()
            escalated += sum(int(st.stage_evals[k - 1][s]) for s in range(hot + 1, 4))
        if got != v["keep"]:
            bad.append((k, v["alpha"], v["rho"], v["keep"], got))
    assert not bad, bad[:5]
    # threshold vectors cannot all be settled by the first stage
    assert escalated > 0
