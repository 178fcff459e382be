"""GPU parity of the batched ensemble entry point (pac_compute_batch).

Every molecule of a batch must get exactly the lists that a run on that
molecule alone gives: the oracle's (reference row order, ids local to the
molecule) and pac_compute's.  Batches include the edge cases of an
ensemble: an empty molecule, one- and two-ball molecules, molecules split
over several pipeline passes.
"""
import sys

import numpy as np
import pytest

import pac
import oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ensemble(n_mol, nmin, nmax, seed):
    """synthetic ensemble plus an empty, a one-ball and a two-ball molecule"""
    xyzr, off = pac.synth_batch(n_mol, nmin, nmax, seed)
    mols = [xyzr[off[m]:off[m + 1]] for m in range(n_mol)]
    extra_one = np.array([[1.0, 2.0, 3.0, 1.7]], dtype=np.float32)
    extra_two = np.array([[0.0, 0.0, 0.0, 1.7], [2.5, 0.0, 0.0, 1.55]],
                         dtype=np.float32)
    empty = np.zeros((0, 4), dtype=np.float32)
    mols = mols[:2] + [empty, extra_one] + mols[2:-1] + [extra_two] + mols[-1:]
    off = np.zeros(len(mols) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(m) for m in mols])
    return np.concatenate(mols).astype(np.float32), off, mols


def _check_mol(br, m, expect, what):
    got = br.lists(m)
    for d in (1, 2, 3):
        assert got[d].shape == expect[d].shape, "%s mol %d dim %d: %s vs %s" % (
            what, m, d, got[d].shape, expect[d].shape)
        assert np.array_equal(got[d], expect[d]), "%s mol %d dim %d differs" % (
            what, m, d)


@pytest.fixture(scope="module")
def ensemble():
    return _ensemble(14, 200, 3000, seed=3)


@pytest.mark.parametrize("prec", [pac.PREC_PLAIN_FP64, pac.PREC_STAGED])
def test_batch_vs_oracle(ensemble, prec):
    xyzr, off, mols = ensemble
    _say("case: batch of %d molecules (%d balls) prec=%d vs oracle"
         % (len(mols), len(xyzr), prec))
    br = pac.compute_batch(xyzr, off, 1.4, 0.0, precision=prec)
    c = br.counts()
    assert list(c[:, 0]) == [len(m) for m in mols]
    st = br.stats()
    assert st.n == len(xyzr)
    for m, mb in enumerate(mols):
        o, _, _ = oracle.run(mb, 1.4, 0.0)
        _check_mol(br, m, o, "batch")
    assert st.edges_final == c[:, 1].sum()
    assert st.tets_final == c[:, 3].sum()


def test_batch_passes_and_single_runs(ensemble):
    """several pipeline passes (batch_balls small) == one pass == per-molecule
    pac_compute, for both cell policies"""
    xyzr, off, mols = ensemble
    for policy in (pac.CELL_EXACT, pac.CELL_REFERENCE_RAW):
        _say("case: batch passes policy=%d" % policy)
        one = pac.compute_batch(xyzr, off, 1.4, 0.0, cell_policy=policy)
        many = pac.compute_batch(xyzr, off, 1.4, 0.0, cell_policy=policy,
                                 batch_balls=2500)
        for m, mb in enumerate(mols):
            single = pac.compute(mb, 1.4, 0.0, cell_policy=policy).lists()
            _check_mol(one, m, single, "one pass")
            _check_mol(many, m, single, "many passes")


def test_batch_canonical_and_alpha(ensemble):
    xyzr, off, mols = ensemble
    _say("case: batch canonical order, alpha 1.5, solvent 0")
    br = pac.compute_batch(xyzr, off, 0.0, 1.5, order=pac.ORDER_CANONICAL)
    for m, mb in enumerate(mols):
        single = pac.compute(mb, 0.0, 1.5, order=pac.ORDER_CANONICAL).lists()
        _check_mol(br, m, single, "canonical")


def test_batch_device_inputs(ensemble):
    import torch
    xyzr, off, mols = ensemble
    _say("case: batch from device memory")
    t = torch.from_numpy(xyzr).to("cuda:0")
    torch.cuda.synchronize()
    br = pac.compute_batch(t.data_ptr(), off, 1.4, 0.0, device_inputs=True)
    host = pac.compute_batch(xyzr, off, 1.4, 0.0)
    for m in range(len(mols)):
        _check_mol(br, m, host.lists(m), "device inputs")
