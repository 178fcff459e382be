"""GPU parity of the FP64-ball pipeline (pac_compute_f64) against the oracle's
FP64 run: continuous coordinates, no rounding, same exact semantics.  FP64
balls are keyed by their FP32 roundings in both, so rows also come out in
the same order.  An FP64 input whose values are all FP32 numbers must give
exactly the FP32 pipeline's lists."""
import sys

import numpy as np
import pytest

import pac
import oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(g, o, what):
    for d in (1, 2, 3):
        assert g[d].shape == o[d].shape, "%s dim %d: %s vs %s" % (
            what, d, g[d].shape, o[d].shape)
        assert np.array_equal(g[d], o[d]), "%s dim %d differs" % (what, d)


def test_c4_20k_f64_vs_oracle():
    _say("case: C4 20k FP64 jittered globule vs oracle")
    b = pac.synth_globule_f64(20000, 45.7, 1.3, seed=4, sigma=0.01)
    assert b.dtype == np.float64
    r = pac.compute_f64(b, 1.4, 0.0)
    o, ost, _ = oracle.run(b, 1.4, 0.0)
    _same(r.lists(), o, "C4")
    st = r.stats()
    assert st.scale_bits == ost["scale"]
    assert st.edge_candidates == ost["E_cand"]
    # the FP64 ladder starts at FP64: nothing is booked at the FP32 stage
    assert all(st.stage_evals[f][0] == 0 for f in range(6))


def test_f64_of_fp32_values_equals_f32_pipeline():
    _say("case: FP64 copy of FP32 balls == FP32 pipeline")
    b = pac.synth_globule(3000, None, 1.3, seed=9)
    # solvent 0 so that both compute the same radii (fl32(r + 1.4f) is not
    # r + 1.4 in FP64); the same grid, keys and exact decisions follow
    for alpha in (2.0, 4.5):
        r32 = pac.compute(b, 0.0, alpha).lists()
        r64 = pac.compute_f64(b.astype(np.float64), 0.0, alpha).lists()
        _same(r64, r32, "f64 of f32 values, alpha %g" % alpha)


def test_f64_lattice_ties_and_canonical():
    _say("case: FP64 lattice (exact ties, SoS) vs oracle, canonical order")
    g = np.arange(5, dtype=np.float64) * 2.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    b = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.full(X.size, 1.7)], 1)
    r = pac.compute_f64(b, 1.4, 0.0)
    o, ost, _ = oracle.run(b, 1.4, 0.0)
    _same(r.lists(), o, "lattice")
    assert r.stats().sos_calls > 0
    rc = pac.compute_f64(b, 1.4, 0.0, order=pac.ORDER_CANONICAL).lists()
    for d in (1, 2, 3):
        srt = np.sort(o[d], axis=1)
        srt = srt[np.lexsort(srt.T[::-1])]
        assert np.array_equal(rc[d], srt)


def test_f64_tiny_and_bad_inputs():
    _say("case: FP64 tiny / invalid inputs")
    assert list(pac.compute_f64(np.zeros((0, 4)), 1.4, 0.0).counts()) == [0, 0, 0, 0]
    two = np.array([[0.1, 0.2, 0.3, 1.7], [2.0, 0.0, 0.0, 1.55]])
    o, _, _ = oracle.run(two, 1.4, 0.0)
    _same(pac.compute_f64(two, 1.4, 0.0).lists(), o, "two balls")
    bad = two.copy()
    bad[1, 0] = np.nan
    with pytest.raises(pac.PacError) as e:
        pac.compute_f64(bad, 1.4, 0.0)
    assert e.value.code == pac.PAC_ERR_INVALID
