"""GPU: a batch recorded as one HIP graph (pac_graph_*).

A graph launch must give exactly the lists of the eager batch
(pac_compute_batch, itself checked against the oracle molecule by
molecule): on replays with eager work and synchronisations in between,
after the input is replaced in place (host copy, and the caller's device
array), and over several pipeline passes.  An input that moves a
molecule's bounds must be reported stale, never returned.
"""
import sys

import numpy as np
import pytest

import pac
from test_gpu_batch import _ensemble, _check_mol

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same_batch(got, want, n_mol, what):
    assert np.array_equal(got.counts(), want.counts()), what
    for m in range(n_mol):
        _check_mol(got, m, want.lists(m), what)


def _swap_centres(xyzr, off, m):
    """swap the centres of two balls of molecule m whose radii differ:
    every coordinate multiset, hence every bound, stays the same while the
    complex changes"""
    x = xyzr.copy()
    a = int(off[m])
    r = x[a:off[m + 1], 3]
    b = a + int(np.nonzero(r != r[0])[0][0])
    x[[a, b], :3] = x[[b, a], :3]
    return x


@pytest.fixture(scope="module")
def ensemble():
    return _ensemble(14, 200, 3000, seed=3)


@pytest.mark.parametrize("prec", [pac.PREC_PLAIN_FP64, pac.PREC_STAGED])
def test_graph_replays_equal_eager(ensemble, prec):
    xyzr, off, mols = ensemble
    _say("case: graph of %d molecules, 4 passes, prec=%d" % (len(mols), prec))
    kw = dict(precision=prec, batch_balls=5000)
    eager = pac.compute_batch(xyzr, off, 1.4, 0.0, **kw)
    g = pac.Graph(xyzr, off, 1.4, 0.0, **kw)
    nodes, nbytes = g.info()
    assert nodes > 50 and nbytes > 0
    est = eager.stats()
    for rep in range(3):
        g.launch()
        r = g.result()
        _same_batch(r, eager, len(mols), "replay %d" % rep)
        st = r.stats()
        for f in ("n", "edge_candidates", "edge_alpha", "tri_candidates",
                  "tri_alpha", "tet_candidates", "tet_alpha", "edges_final",
                  "tris_final", "tets_final"):
            assert getattr(st, f) == getattr(est, f), f
        assert st.total_ms > 0
        pac.compute(mols[0], 1.4, 0.0)  # eager work between replays


def test_graph_new_input_and_stale(ensemble):
    xyzr, off, mols = ensemble
    _say("case: graph input replaced in place, then made stale")
    g = pac.Graph(xyzr, off, 1.4, 0.0, batch_balls=5000)
    x2 = _swap_centres(xyzr, off, 4)
    x2 = _swap_centres(x2, off, len(mols) - 1)
    g.set_input(x2)
    g.launch()
    _same_batch(g.result(), pac.compute_batch(x2, off, 1.4, 0.0), len(mols),
                "new input")
    x3 = x2.copy()
    x3[off[5]:off[6], :3] *= np.float32(1.25)
    g.set_input(x3)
    g.launch()
    with pytest.raises(pac.PacError) as e:
        g.result()
    assert e.value.code == pac.PAC_ERR_STALE
    g.set_input(xyzr)  # back to the recorded geometry
    g.launch()
    _same_batch(g.result(), pac.compute_batch(xyzr, off, 1.4, 0.0), len(mols),
                "restored input")


def test_graph_device_inputs_single_pass(ensemble):
    import torch
    xyzr, off, mols = ensemble
    _say("case: graph over the caller's device array, one pass, alpha 1.5")
    t = torch.from_numpy(xyzr).to("cuda:0")
    torch.cuda.synchronize()
    g = pac.Graph(t.data_ptr(), off, 0.0, 1.5, device_inputs=True)
    g.launch()
    _same_batch(g.result(), pac.compute_batch(xyzr, off, 0.0, 1.5), len(mols),
                "device inputs")
    x2 = _swap_centres(xyzr, off, 7)
    t.copy_(torch.from_numpy(x2))
    torch.cuda.synchronize()
    g.launch()
    _same_batch(g.result(), pac.compute_batch(x2, off, 0.0, 1.5), len(mols),
                "device inputs, new values")


def test_graph_raw_policy_and_skip(ensemble):
    xyzr, off, mols = ensemble
    _say("case: graph with the reference's raw cell policy / skipped tet test "
         "/ quick rejection pass")
    for kw in (dict(cell_policy=pac.CELL_REFERENCE_RAW), dict(skip_tet_ac2=True),
               dict(ac2_fine=4, ac2_quick=1)):
        g = pac.Graph(xyzr, off, 1.4, 0.0, batch_balls=7000, **kw)
        g.launch()
        _same_batch(g.result(), pac.compute_batch(xyzr, off, 1.4, 0.0, **kw),
                    len(mols), str(kw))
