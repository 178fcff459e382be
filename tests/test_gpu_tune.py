"""GPU: the autotuner picks settings from its candidate set, and the
settings never change the output (lists identical to the defaults)."""
import sys

import numpy as np
import pytest

import pac

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_autotune_single_and_batch():
    _say("case: autotune 20k globule and a small batch")
    b = pac.synth_globule(20000, None, 1.3, seed=12)
    tuned, ms = pac.autotune(b, 1.4, 0.0)
    assert tuned["ac2_budget"] in (4, 6, 8, 12, 16, 24)
    assert tuned["ac2_grid"] in (1024, 2048, 4096)
    assert ms > 0
    ref = pac.compute(b, 1.4, 0.0).lists()
    got = pac.compute(b, 1.4, 0.0, **tuned).lists()
    for d in (1, 2, 3):
        assert np.array_equal(ref[d], got[d])
    x, off = pac.synth_batch(12, 200, 4000, seed=13)
    tb, _ = pac.autotune(x, 1.4, 0.0, offsets=off)
    r1 = pac.compute_batch(x, off, 1.4, 0.0)
    r2 = pac.compute_batch(x, off, 1.4, 0.0, **tb)
    for m in range(12):
        a, c = r1.lists(m), r2.lists(m)
        for d in (1, 2, 3):
            assert np.array_equal(a[d], c[d])


@pytest.mark.parametrize("budget,grid", [(1, 64), (0, 0), (-1, 0), (64, 65536)])
def test_extreme_settings_same_output(budget, grid):
    _say("case: budget %d grid %d" % (budget, grid))
    b = pac.synth_globule(5000, None, 1.3, seed=14)
    ref = pac.compute(b, 1.4, 0.0).lists()
    got = pac.compute(b, 1.4, 0.0, ac2_budget=budget, ac2_grid=grid).lists()
    for d in (1, 2, 3):
        assert np.array_equal(ref[d], got[d])
