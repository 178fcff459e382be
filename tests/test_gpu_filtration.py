"""GPU: the filtration of the weighted alpha complexes
(pac_options.filtration, pac_result_copy_birth / _filtration).

What fixes the values: the complex at any alpha' <= alpha must be exactly
the simplices of the complex at alpha whose value is <= alpha'.  That is
checked against direct pipeline runs at several alpha' (each itself checked
against the oracle for one alpha'), leaving out only simplices whose FP64
value lies within 1e-9 of alpha'.  Also: values never decrease from a face
to a coface, an edge / triangle that is no face of the next dimension is
at its own rho, one that is not at its own rho is at the smallest value of
its cofaces, balls at -r^2, and the sorted filtration lists every simplex
once, after its faces.
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


def _keyset(rows):
    return set(map(tuple, np.sort(rows, axis=1).tolist()))


def _faces(rows):
    k = rows.shape[1]
    out = []
    for drop in range(k):
        cols = [c for c in range(k) if c != drop]
        out.append(np.sort(rows[:, cols], axis=1))
    return out


def _index(rows):
    return {t: i for i, t in enumerate(map(tuple, np.sort(rows, axis=1).tolist()))}


def _check_subcomplexes(b, solvent, alpha, r, betas, what, check_oracle=None):
    L = r.lists()
    B = {d: r.births(d) for d in (1, 2, 3)}
    for beta in betas:
        lb = pac.compute(b, solvent, beta).lists()
        for d in (1, 2, 3):
            full = np.sort(L[d], axis=1)
            tol = 1e-9 * max(1.0, abs(beta))
            near = np.abs(B[d] - beta) <= tol
            want = _keyset(full[(B[d] <= beta) & ~near])
            got = _keyset(lb[d]) - _keyset(full[near])
            assert got == want, "%s beta=%g dim %d: %d extra, %d missing" % (
                what, beta, d, len(got - want), len(want - got))
        if check_oracle is not None and beta == check_oracle:
            o, _, _ = oracle.run(b, solvent, beta)
            for d in (1, 2, 3):
                assert _keyset(o[d]) == _keyset(lb[d]), "oracle beta=%g" % beta


def _check_structure(b, solvent, alpha, r, what):
    L = r.lists()
    B = {d: r.births(d) for d in (0, 1, 2, 3)}
    rho = {d: r.alpha_values(d) for d in (1, 2, 3)}
    rad = b[:, 3].astype(np.float32) + np.float32(solvent)
    assert np.array_equal(B[0], -(rad.astype(np.float64) ** 2)), what
    assert np.array_equal(B[3], rho[3]), what
    for d in (1, 2, 3):
        assert np.all(B[d] <= alpha + 1e-9 * max(1, abs(alpha))), (what, d)
    for d in (1, 2):
        idx = _index(L[d])
        cof = L[d + 1]
        low = np.full(len(L[d]), np.inf)
        for f in _faces(cof):
            rows = np.array([idx[t] for t in map(tuple, f.tolist())], np.int64)
            # a face is never later than its coface
            assert np.all(B[d][rows] <= B[d + 1]), (what, d)
            np.minimum.at(low, rows, B[d + 1])
        own = B[d] == rho[d]
        # no coface: own rho; otherwise own rho or the smallest coface value
        assert np.all(own[np.isinf(low)]), (what, d)
        assert np.all(own | (B[d] == low)), (what, d)
        assert np.all(B[d] <= low), (what, d)
    # edges from balls
    assert np.all(B[0][L[1][:, 0]] <= B[1]) and np.all(B[0][L[1][:, 1]] <= B[1])


def _check_order(r, what):
    c = [int(x) for x in r.counts()]
    d, i, v = r.filtration()
    assert len(d) == sum(c)
    assert np.all(np.diff(v) >= 0), what
    pos = {}
    for q in range(4):
        sel = np.nonzero(d == q)[0]
        assert sorted(i[sel].tolist()) == list(range(c[q])), (what, q)
        p = np.empty(c[q], np.int64)
        p[i[sel]] = sel
        pos[q] = p
    Bs = {q: r.births(q) for q in range(4)}
    for q in range(4):
        assert np.array_equal(v[pos[q]], Bs[q]), (what, q)
    L = r.lists()
    for q in (1, 2, 3):
        if q == 1:
            for col in (0, 1):
                assert np.all(pos[0][L[1][:, col]] < pos[1]), what
            continue
        idx = _index(L[q - 1])
        for f in _faces(L[q]):
            rows = np.array([idx[t] for t in map(tuple, f.tolist())], np.int64)
            assert np.all(pos[q - 1][rows] < pos[q]), (what, q)


def test_filtration_globule():
    _say("case: filtration, 3000-ball globule, solvent 1.4, alpha 2.0")
    b = pac.synth_globule(3000, None, 1.3, seed=21)
    r = pac.compute(b, 1.4, 2.0, filtration=True)
    _check_structure(b, 1.4, 2.0, r, "globule")
    _check_order(r, "globule")
    _check_subcomplexes(b, 1.4, 2.0, r, [-1.5, -0.75, 0.0, 0.6, 1.3],
                        "globule", check_oracle=0.6)
    # the lists themselves are those of a run without the filtration
    plain = pac.compute(b, 1.4, 2.0).lists()
    for d in (1, 2, 3):
        assert np.array_equal(r.lists()[d], plain[d])


def test_filtration_solvent0_and_plain_ladder():
    _say("case: filtration, 2000 balls, solvent 0, alpha 3.0, plain FP64")
    b = pac.synth_globule(2000, None, 1.3, seed=5)
    r = pac.compute(b, 0.0, 3.0, filtration=True, precision=pac.PREC_PLAIN_FP64)
    _check_structure(b, 0.0, 3.0, r, "s0")
    _check_subcomplexes(b, 0.0, 3.0, r, [-2.0, 0.0, 1.0, 2.2], "s0")


def test_filtration_lattice_ties():
    """Equal radii on a cubic lattice: many simplices share a value, ties
    of the emptiness tests everywhere."""
    _say("case: filtration, 6^3 lattice, alpha 1.5")
    m = 6
    g = np.stack(np.meshgrid(*[np.arange(m) * 2.0] * 3, indexing="ij"), -1).reshape(-1, 3)
    b = np.concatenate([g, np.full((len(g), 1), 1.7)], 1).astype(np.float32)
    r = pac.compute(b, 0.0, 1.5, filtration=True)
    _check_structure(b, 0.0, 1.5, r, "lattice")
    _check_order(r, "lattice")
    vals = np.unique(np.concatenate([r.births(d) for d in (1, 2, 3)]))
    mids = [(vals[k] + vals[k + 1]) / 2 for k in range(0, len(vals) - 1,
                                                       max(1, len(vals) // 4))]
    _check_subcomplexes(b, 0.0, 1.5, r, mids[:4], "lattice")


def test_filtration_chunked_equals_one_pass():
    _say("case: filtration through the chunked driver, 12k balls, chunk 2048")
    b = pac.synth_globule(12000, None, 1.3, seed=8)
    one = pac.compute(b, 1.4, 1.0, filtration=True)
    ch = pac.compute(b, 1.4, 1.0, filtration=True, chunk_balls=2048)
    for d in (0, 1, 2, 3):
        assert np.array_equal(ch.births(d), one.births(d)), d
    for x, y in zip(ch.filtration(), one.filtration()):
        assert np.array_equal(x, y)


def test_filtration_argument_errors():
    _say("case: filtration argument errors")
    b = pac.synth_globule(300, None, 1.3, seed=2)
    r = pac.compute(b, 1.4, 0.0)
    with pytest.raises(pac.PacError) as e:
        r.births(2)
    assert e.value.code == pac.PAC_ERR_INVALID
    with pytest.raises(pac.PacError) as e:
        pac.compute(b, 1.4, 0.0, filtration=True, order=pac.ORDER_CANONICAL)
    assert e.value.code == pac.PAC_ERR_INVALID
    with pytest.raises(pac.PacError) as e:
        pac.compute_batch(b, np.array([0, 150, 300]), 1.4, 0.0, filtration=True)
    assert e.value.code == pac.PAC_ERR_INVALID
    with pytest.raises(pac.PacError) as e:
        pac.compute(b, 1.4, 0.0, filtration=True, skip_tet_ac2=True)
    assert e.value.code == pac.PAC_ERR_INVALID
    r0 = pac.compute(np.zeros((0, 4), np.float32), 1.4, 0.0, filtration=True)
    d, i, v = r0.filtration()
    assert len(d) == 0
