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
    assert tuned["ac2_budget"] in (-1, 2, 4, 8, 16)
    assert tuned["ac2_grid"] in (1024, 2048, 4096)
    assert tuned["ac2_fine"] in (-1, 2, 4)
    assert tuned["ac2_quick"] in (0, 1)
    if tuned["ac2_budget"] > 0:  # a second phase: its shape is tuned too
        assert tuned["block"] in (64, 128, 256)
        assert tuned["ac2_chunk"] in (64, 256, 1024)
    else:  # one phase: the given (default) shape comes back
        assert tuned["block"] == 0 and tuned["ac2_chunk"] == 0
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


@pytest.mark.parametrize("fine,quick", [(-1, 0), (1, 0), (2, 0), (3, 0), (4, 0),
                                        (8, 0), (2, 1), (4, 1)])
def test_scan_grids_same_output(fine, quick):
    """the emptiness scans give the same lists over any scan grid (the
    reference cells or cells split along x) with or without the quick
    rejection pass; on a globule, a degenerate lattice (exact ties, SoS)
    and a small batch"""
    _say("case: scan grid %d quick %d" % (fine, quick))
    b = pac.synth_globule(5000, None, 1.3, seed=15)
    m = 6
    g = np.stack(np.meshgrid(*[np.arange(m) * 2.0] * 3, indexing="ij"), -1).reshape(-1, 3)
    lat = np.concatenate([g, np.full((len(g), 1), 1.7)], 1).astype(np.float32)
    for what, x in (("globule", b), ("lattice", lat)):
        for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
            ref = pac.compute(x, 1.4, 0.0, precision=prec, ac2_fine=-1).lists()
            got = pac.compute(x, 1.4, 0.0, precision=prec, ac2_fine=fine,
                              ac2_quick=quick).lists()
            for d in (1, 2, 3):
                assert np.array_equal(ref[d], got[d]), (what, prec, d)
    xb, off = pac.synth_batch(9, 200, 3000, seed=16)
    r1 = pac.compute_batch(xb, off, 1.4, 0.0, ac2_fine=-1)
    r2 = pac.compute_batch(xb, off, 1.4, 0.0, ac2_fine=fine, ac2_quick=quick,
                           batch_balls=4000)
    for k in range(9):
        a, c = r1.lists(k), r2.lists(k)
        for d in (1, 2, 3):
            assert np.array_equal(a[d], c[d]), ("batch", k, d)


@pytest.mark.parametrize("block,chunk", [(64, 64), (128, 1000), (256, 64),
                                         (256, 4096)])
def test_persistent_phase_shapes_same_output(block, chunk):
    """the persistent second emptiness phase gives the same lists for any
    block shape and queue granularity (a chunk not a multiple of 64 is
    rounded up), on a globule with a small budget (much work left for the
    second phase), the degenerate lattice, and a batch recorded as one
    graph"""
    _say("case: persistent phase block %d chunk %d" % (block, chunk))
    b = pac.synth_globule(6000, None, 1.3, seed=17)
    m = 6
    g = np.stack(np.meshgrid(*[np.arange(m) * 2.0] * 3, indexing="ij"), -1).reshape(-1, 3)
    lat = np.concatenate([g, np.full((len(g), 1), 1.7)], 1).astype(np.float32)
    for what, x in (("globule", b), ("lattice", lat)):
        for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
            ref = pac.compute(x, 1.4, 0.0, precision=prec).lists()
            got = pac.compute(x, 1.4, 0.0, precision=prec, ac2_budget=2,
                              ac2_grid=96, block=block, ac2_chunk=chunk).lists()
            for d in (1, 2, 3):
                assert np.array_equal(ref[d], got[d]), (what, prec, d)
    xb, off = pac.synth_batch(7, 200, 3000, seed=18)
    r1 = pac.compute_batch(xb, off, 1.4, 0.0)
    gr = pac.Graph(xb, off, 1.4, 0.0, ac2_budget=2, ac2_grid=96, block=block,
                   ac2_chunk=chunk)
    for _ in range(2):
        gr.launch()
        r2 = gr.result()
        for k in range(7):
            a, c = r1.lists(k), r2.lists(k)
            for d in (1, 2, 3):
                assert np.array_equal(a[d], c[d]), ("graph", k, d)
