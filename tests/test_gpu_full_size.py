"""GPU parity at the full sizes of the benchmark workloads.

* 100k-ball globule (the headline workload, same generator and seed as the
  benchmark): both precision ladders against the oracle, row for row.
* 2,000,000-ball degenerate stress case in one pass: the oracle would take
  minutes, so size-independent properties stand in for it -- the staged and
  the plain FP64 ladders (independent filter arithmetic down to the same
  exact integers) give the same lists, the chunked driver (a different
  decomposition of the same complex) gives the same lists, the exact stages
  and SoS are reached, and the complex is closed (every face of a listed
  simplex is listed).
* 4096-molecule ensemble as one graph launch: staged equals plain for every
  molecule, a sample equals single-molecule runs and the smallest molecules
  of that sample equal the oracle.
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


def _threads():
    return min(16, os.cpu_count() or 1)


def _keys(rows, n):
    """one int64 key per row of sorted vertex ids (n^3 < 2^63 for n <= 2M)"""
    s = np.sort(rows.astype(np.int64), axis=1)
    k = s[:, 0]
    for c in range(1, s.shape[1]):
        k = k * n + s[:, c]
    return k


def _closed(L, n):
    """every face of a listed tet / triangle is a listed triangle / edge"""
    for d in (3, 2):
        have = np.sort(_keys(L[d - 1], n))
        cols = range(d + 1)
        for drop in cols:
            f = _keys(L[d][:, [c for c in cols if c != drop]], n)
            at = np.searchsorted(have, f)
            at[at == len(have)] = 0
            if not np.all(have[at] == f):
                return False
    return bool(np.all(L[1] >= 0) and np.all(L[1] < n))


def test_globule_100k_vs_oracle():
    _say("case: 100k globule (radius 78.2, seed 2, 0.001 A) vs oracle")
    b = pac.synth_globule(100000, 78.2, 1.3, seed=2, decimals=3)
    o, ost, _ = oracle.run(b, 1.4, 0.0, nthreads=_threads())
    for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
        r = pac.compute(b, 1.4, 0.0, precision=prec)
        L, st = r.lists(), r.stats()
        for d in (1, 2, 3):
            assert np.array_equal(L[d], o[d]), "prec %d dim %d" % (prec, d)
        assert st.edge_candidates == ost["E_cand"]
        assert st.tet_alpha == ost["K_alpha"]
        hot = 0 if prec == pac.PREC_STAGED else 1
        # every candidate of each filter is classified at the first stage
        assert st.stage_evals[2][hot] == st.tet_candidates
        assert st.stage_evals[3][hot] == st.ac2_pairs[0]


def test_stress_2m_one_pass_properties():
    _say("case: 2M stress (lattice block 10%) one pass, staged vs plain vs chunked")
    b = pac.synth_stress(2000000, 212.2, seed=5, lattice_frac=0.1, spacing=2.0,
                         lattice_r=1.70)
    rs = pac.compute(b, 1.4, 0.0, precision=pac.PREC_STAGED)
    Ls, st = rs.lists(), rs.stats()
    del rs
    assert st.stage_evals[3][3] > 0, "tet emptiness never reached wide integers"
    assert st.sos_calls > 0
    _say("  staged: counts %s" % [len(Ls[d]) for d in (1, 2, 3)])
    rp = pac.compute(b, 1.4, 0.0, precision=pac.PREC_PLAIN_FP64)
    Lp = rp.lists()
    del rp
    for d in (1, 2, 3):
        assert np.array_equal(Ls[d], Lp[d]), "staged vs plain dim %d" % d
    del Lp
    _say("  plain FP64 ladder: same lists")
    rc = pac.compute(b, 1.4, 0.0, chunk_balls=262144)
    Lc = rc.lists()
    del rc
    for d in (1, 2, 3):
        assert np.array_equal(Ls[d], Lc[d]), "one pass vs chunked dim %d" % d
    del Lc
    _say("  chunked driver: same lists; checking closure")
    assert _closed(Ls, len(b))


def test_batch_4096_graph():
    _say("case: 4096-molecule ensemble, one graph launch, staged vs plain")
    xyzr, off = pac.synth_batch(4096, 200, 20000, seed=3)
    res = {}
    for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
        g = pac.Graph(xyzr, off, 1.4, 0.0, precision=prec)
        g.launch()
        res[prec] = g.result()
        del g
    bs, bp = res[pac.PREC_STAGED], res[pac.PREC_PLAIN_FP64]
    cs, cp = bs.counts(), bp.counts()
    assert np.array_equal(cs, cp)
    assert list(cs[:, 0]) == list(np.diff(off))
    for m in range(4096):
        ls, lp = bs.lists(m), bp.lists(m)
        for d in (1, 2, 3):
            assert np.array_equal(ls[d], lp[d]), "mol %d dim %d" % (m, d)
    rng = np.random.default_rng(11)
    sample = sorted(rng.choice(4096, 16, replace=False).tolist())
    for m in sample:
        mb = xyzr[off[m]:off[m + 1]]
        single = pac.compute(mb, 1.4, 0.0).lists()
        got = bs.lists(m)
        for d in (1, 2, 3):
            assert np.array_equal(got[d], single[d]), "single mol %d dim %d" % (m, d)
    for m in sorted(sample, key=lambda q: off[q + 1] - off[q])[:4]:
        o, _, _ = oracle.run(xyzr[off[m]:off[m + 1]], 1.4, 0.0, nthreads=_threads())
        got = bs.lists(m)
        for d in (1, 2, 3):
            assert np.array_equal(got[d], o[d]), "oracle mol %d dim %d" % (m, d)
