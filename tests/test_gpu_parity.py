"""GPU parity: the HIP pipeline against the CPU oracle (same inputs, same
exact semantics) and against the reference's own output fixture.

Integer outputs (simplex lists) must match bit for bit, including row
order.  The reference's fixture output_files/1grm_pAC_LD_alpha1.txt
(solvent 0, alpha 1, tet emptiness effectively skipped when it was made) is
reproduced byte for byte through pac_write_output.
"""
import os
import sys

import numpy as np
import pytest

import pac
import oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gpu_lists(b, s, a, **kw):
    r = pac.compute(b, s, a, **kw)
    return r.lists(), r.stats()


def _assert_same(g, o, what):
    for d in (1, 2, 3):
        assert g[d].shape == o[d].shape, "%s dim %d: gpu %s oracle %s" % (
            what, d, g[d].shape, o[d].shape)
        assert np.array_equal(g[d], o[d]), "%s dim %d differs" % (what, d)


@pytest.fixture(scope="module")
def crd_1grm(golden_dir):
    return pac.read_crd(os.path.join(golden_dir, "1grm.crd"))


@pytest.mark.parametrize("prec", [pac.PREC_PLAIN_FP64, pac.PREC_STAGED])
def test_1grm_fixture_bytes(crd_1grm, golden_dir, tmp_path, prec):
    _say("case: 1grm fixture skip_tet_ac2 prec=%d" % prec)
    r = pac.compute(crd_1grm, 0.0, 1.0, skip_tet_ac2=True, precision=prec)
    assert list(r.counts()) == [260, 842, 724, 183]
    out = tmp_path / "out.txt"
    pac.write_output(str(out), crd_1grm, 0.0, r)
    ref = open(os.path.join(golden_dir, "1grm_pAC_LD_alpha1.txt"), "rb").read()
    assert out.read_bytes() == ref


@pytest.mark.parametrize("prec", [pac.PREC_PLAIN_FP64, pac.PREC_STAGED])
@pytest.mark.parametrize("policy", [pac.CELL_EXACT, pac.CELL_REFERENCE_RAW])
def test_1grm_full(crd_1grm, prec, policy):
    _say("case: 1grm full prec=%d policy=%d" % (prec, policy))
    g, st = _gpu_lists(crd_1grm, 0.0, 1.0, precision=prec, cell_policy=policy)
    o, ost, _ = oracle.run(crd_1grm, 0.0, 1.0, cell_policy=policy)
    _assert_same(g, o, "1grm")
    assert [len(g[d]) for d in (1, 2, 3)] == [812, 650, 136]


@pytest.fixture(scope="module")
def crd_1k4c(golden_dir):
    return pac.read_crd(os.path.join(golden_dir, "1k4c.crd"))


@pytest.mark.parametrize("s,a", [(1.4, 0.0), (0.0, 1.0)])
def test_1k4c_full(crd_1k4c, s, a):
    """The reference's own medium protein (data/1k4c.crd, 16,068 atoms with
    the four CRD radii), full pipeline, both precisions and cell policies."""
    _say("case: 1k4c full s=%g a=%g" % (s, a))
    for policy in (pac.CELL_EXACT, pac.CELL_REFERENCE_RAW):
        o, ost, _ = oracle.run(crd_1k4c, s, a, cell_policy=policy)
        for prec in (pac.PREC_PLAIN_FP64, pac.PREC_STAGED):
            g, st = _gpu_lists(crd_1k4c, s, a, precision=prec, cell_policy=policy)
            _assert_same(g, o, "1k4c s=%g a=%g policy=%d prec=%d" % (s, a, policy, prec))
            assert st.edge_candidates == ost["E_cand"]


@pytest.mark.parametrize("prec", [pac.PREC_PLAIN_FP64, pac.PREC_STAGED])
def test_c1_1000(prec):
    _say("case: C1 1000 balls prec=%d" % prec)
    b = pac.synth_globule(1000, 16.84, 1.3, seed=1)
    g, st = _gpu_lists(b, 1.4, 0.0, precision=prec)
    o, ost, _ = oracle.run(b, 1.4, 0.0)
    _assert_same(g, o, "C1")
    assert st.edge_candidates == ost["E_cand"]
    assert st.tri_candidates == ost["T_cand"]
    assert st.tet_candidates == ost["K_cand"]
    assert st.tet_alpha == ost["K_alpha"]


def test_c1_solvent0_alpha_positive():
    _say("case: C1 1000 balls s=0 alpha=2")
    b = pac.synth_globule(1000, 16.84, 1.3, seed=11)
    g, st = _gpu_lists(b, 0.0, 2.0)
    o, ost, _ = oracle.run(b, 0.0, 2.0)
    _assert_same(g, o, "C1 s0 a2")


def test_medium_20k_staged():
    _say("case: 20k balls staged")
    b = pac.synth_globule(20000, None, 1.3, seed=7)
    g, st = _gpu_lists(b, 1.4, 0.0, precision=pac.PREC_STAGED)
    o, ost, _ = oracle.run(b, 1.4, 0.0)
    _assert_same(g, o, "20k")


def test_lattice_degenerate():
    """Cubic lattice with equal radii: every cube is co-spherical, so the
    emptiness test runs into exact ties and SoS."""
    _say("case: 6^3 lattice degenerate")
    m = 6
    g = np.stack(np.meshgrid(*[np.arange(m) * 2.0] * 3, indexing="ij"), -1).reshape(-1, 3)
    b = np.concatenate([g, np.full((len(g), 1), 1.7)], 1).astype(np.float32)
    for prec in (pac.PREC_PLAIN_FP64, pac.PREC_STAGED):
        gl, st = _gpu_lists(b, 1.4, 0.0, precision=prec)
        o, ost, _ = oracle.run(b, 1.4, 0.0)
        _assert_same(gl, o, "lattice")
        assert st.sos_calls > 0


def test_stress_lattice_block():
    """The degenerate stress workload at 60,000 balls: Poisson-disk globule
    with a central block replaced by a 2.0 A lattice of equal radii (10%).
    Exact ties reach the wide-integer stage and SoS."""
    _say("case: stress 60k with lattice block")
    b = pac.synth_stress(60000, None, seed=5, lattice_frac=0.1, spacing=2.0,
                         lattice_r=1.70)
    o, ost, _ = oracle.run(b, 1.4, 0.0)
    for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
        g, st = _gpu_lists(b, 1.4, 0.0, precision=prec)
        _assert_same(g, o, "stress")
        assert st.sos_calls > 0
        assert st.stage_evals[3][3] > 0  # tet emptiness reached wide integers


def test_empty_and_tiny():
    _say("case: empty / tiny inputs")
    r = pac.compute(np.zeros((0, 4), np.float32), 1.4, 0.0)
    assert list(r.counts()) == [0, 0, 0, 0]
    r = pac.compute(np.array([[0, 0, 0, 1.5]], np.float32), 1.4, 0.0)
    assert list(r.counts()) == [1, 0, 0, 0]
    b = np.array([[0, 0, 0, 1.5], [1, 0, 0, 1.5]], np.float32)
    g, _ = _gpu_lists(b, 0.0, 0.0)
    assert g[1].tolist() == [[0, 1]]


def test_duplicate_centres():
    _say("case: duplicate centres")
    b = pac.synth_globule(300, None, 1.3, seed=5)
    b = np.concatenate([b, b[:20]], 0)
    g, _ = _gpu_lists(b, 1.4, 0.0)
    o, _, _ = oracle.run(b, 1.4, 0.0)
    _assert_same(g, o, "dups")
