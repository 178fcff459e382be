"""CPU: the oracle against the reference's own fixture and against exact
known-answer vectors (tests/golden/predicates.json, made with Fractions by
tests/golden/make_golden.py)."""
import json
import os

import numpy as np
import pytest

import oracle
import pac

LD_COUNTS = (842, 724, 183)       # output_files/1grm_pAC_LD_alpha1.txt
EXACT_1GRM = (812, 650, 136)      # exact weighted alpha complex, s=0, alpha=1


def _crd(golden_dir, name):
    return pac.read_crd(os.path.join(golden_dir, name))


def _write(b, s, out):
    lines = ["%d %d %d %d" % (len(b), len(out[1]), len(out[2]), len(out[3]))]
    for x, y, z, r in b:
        rr = np.float32(np.float32(r) + np.float32(s))
        lines.append("%9.5f %9.5f %9.5f %9.5f" % (float(x), float(y), float(z), float(rr)))
    for d in (1, 2, 3):
        for row in out[d]:
            lines.append(" ".join("%9d" % v for v in row))
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("policy", [0, 1])
def test_ld_fixture_byte_for_byte(golden_dir, policy):
    b = _crd(golden_dir, "1grm.crd")
    out, st, _ = oracle.run(b, 0.0, 1.0, skip_tet_ac2=True, cell_policy=policy)
    assert tuple(len(out[d]) for d in (1, 2, 3)) == LD_COUNTS
    ref = open(os.path.join(golden_dir, "1grm_pAC_LD_alpha1.txt")).read()
    assert _write(b, 0.0, out) == ref


@pytest.mark.parametrize("policy", [0, 1])
def test_1grm_exact_complex(golden_dir, policy):
    b = _crd(golden_dir, "1grm.crd")
    out, st, _ = oracle.run(b, 0.0, 1.0, cell_policy=policy)
    assert tuple(len(out[d]) for d in (1, 2, 3)) == EXACT_1GRM


def test_rows_sorted_in_reference_order(golden_dir):
    """Rows are lexicographic in cell-sorted index space and written as
    original ids (AlphaComplexParallel.cu:1157 sort, :1842-1856 id
    mapping)."""
    b = _crd(golden_dir, "1grm.crd")
    out, _, _ = oracle.run(b, 0.0, 1.0)
    import ctypes
    L = oracle.lib()
    h = ctypes.c_void_p()
    a = np.ascontiguousarray(b, dtype=np.float32)
    assert L.pac_oracle_run_f32(a.ctypes.data, len(a), 0.0, 1.0, 0, 0, 1, ctypes.byref(h)) == 0
    perm = np.zeros(len(a), dtype=np.int32)
    L.pac_oracle_perm(h, perm.ctypes.data)
    L.pac_oracle_free(h)
    rank = np.empty_like(perm)
    rank[perm] = np.arange(len(perm))
    for d in (1, 2, 3):
        s = rank[out[d]]
        assert (np.diff(s, axis=1) > 0).all()
        keys = [tuple(r) for r in s]
        assert keys == sorted(keys)


def _ball(h):
    return [float.fromhex(v) for v in h]


def test_filter_vectors(golden_dir):
    d = json.load(open(os.path.join(golden_dir, "predicates.json")))
    bad = []
    for v in d["filters"]:
        bs = [_ball(b) for b in v["balls"]]
        al = float.fromhex(v["alpha"])
        if len(bs) == 2:
            got = oracle.edge_keep(bs[0], bs[1], al)
        elif len(bs) == 3:
            got = oracle.tri_keep(*bs, al)
        else:
            got = oracle.tet_keep(*bs, al)
        if got != v["keep"]:
            bad.append(v)
    assert not bad, bad[:3]


def test_emptiness_vectors(golden_dir):
    d = json.load(open(os.path.join(golden_dir, "predicates.json")))
    bad = []
    for v in d["empty_low"]:
        bs = [_ball(b) for b in v["balls"]]
        e = _ball(v["e"])
        got = oracle.tri_violated(*bs, e) if len(bs) == 3 else oracle.edge_violated(*bs, e)
        if got != v["violated"]:
            bad.append(v)
    assert not bad, bad[:3]


def test_oracle_deterministic_across_threads():
    b = pac.synth_globule(2000, None, 1.3, seed=9)
    o1, s1, _ = oracle.run(b, 1.4, 0.0, nthreads=1)
    o8, s8, _ = oracle.run(b, 1.4, 0.0, nthreads=8)
    for d in (1, 2, 3):
        assert np.array_equal(o1[d], o8[d])
    assert s1 == s8 or all(s1[k] == s8[k] for k in ("E_cand", "T_cand", "K_cand"))
