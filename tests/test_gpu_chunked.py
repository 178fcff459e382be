"""GPU: the chunked (LD) driver, pac_options.chunk_balls.

The reference's parallelac-largeData walks the cell-sorted balls in chunks
of 512 * 2^k >= <chunk-size> and keeps, per chunk, the simplices whose
lowest sorted vertex lies in it (AlphaComplexLargeData.cu:1227-1233).  With
exact decisions a chunked run must give exactly the lists of one pass: the
same sets, in the same reference row order.  Checked against the oracle
(one pass on the CPU) for chunk sizes from 512 (tens of chunks, most of them
cut through cells) up, for both cell policies and precisions, on a
degenerate lattice block whose ties straddle chunk boundaries, on FP64
inputs, in canonical order, and through the largeData command line.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import pac
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "parallelac-xmp_amd", "bin")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _same(g, o, what):
    for d in (1, 2, 3):
        assert g[d].shape == o[d].shape, "%s dim %d: %s vs %s" % (
            what, d, g[d].shape, o[d].shape)
        assert np.array_equal(g[d], o[d]), "%s dim %d differs" % (what, d)


def _chunks(n, c):
    k = 512
    while k < c:
        k *= 2
    return -(-n // k)


@pytest.fixture(scope="module")
def glob20k():
    b = pac.synth_globule(20000, None, 1.3, seed=7)
    o, ost, _ = oracle.run(b, 1.4, 0.0)
    return b, o


@pytest.mark.parametrize("chunk", [1, 2048, 9000])
def test_chunked_equals_oracle(glob20k, chunk):
    b, o = glob20k
    _say("case: 20k globule, chunk_balls=%d (%d chunks), staged" %
         (chunk, _chunks(len(b), chunk)))
    r = pac.compute(b, 1.4, 0.0, chunk_balls=chunk)
    _same(r.lists(), o, "chunk %d" % chunk)
    st = r.stats()
    assert st.n == len(b)
    assert (st.edges_final, st.tris_final, st.tets_final) == tuple(
        len(o[d]) for d in (1, 2, 3))
    # the halo is computed again by its neighbours
    one = pac.compute(b, 1.4, 0.0).stats()
    assert st.tet_alpha >= one.tet_alpha
    assert st.total_ms > 0 and sum(st.phase_ms) > 0


def test_chunked_plain_and_alpha_values(glob20k):
    b, o = glob20k
    _say("case: 20k globule, chunk 4096, plain FP64 ladder, alpha values")
    r = pac.compute(b, 1.4, 0.0, chunk_balls=4096,
                    precision=pac.PREC_PLAIN_FP64)
    _same(r.lists(), o, "plain chunk 4096")
    one = pac.compute(b, 1.4, 0.0)
    for d in (1, 2, 3):
        assert np.array_equal(r.alpha_values(d), one.alpha_values(d)), d


def test_chunked_reference_raw_policy():
    _say("case: 20k globule, reference raw cells, chunk 2048")
    b = pac.synth_globule(20000, None, 1.3, seed=17)
    o, _, _ = oracle.run(b, 1.4, 0.0, cell_policy=pac.CELL_REFERENCE_RAW)
    r = pac.compute(b, 1.4, 0.0, chunk_balls=2048,
                    cell_policy=pac.CELL_REFERENCE_RAW)
    _same(r.lists(), o, "raw chunk 2048")


def test_chunked_alpha_positive_solvent0():
    _say("case: 8k globule, solvent 0, alpha 2.5, chunk 512")
    b = pac.synth_globule(8000, None, 1.3, seed=23)
    o, _, _ = oracle.run(b, 0.0, 2.5)
    r = pac.compute(b, 0.0, 2.5, chunk_balls=512)
    _same(r.lists(), o, "s0 a2.5 chunk 512")


def test_chunked_stress_lattice_block():
    """Exact ties and SoS in a lattice block that several chunks cut."""
    _say("case: stress 30k with lattice block, chunk 4096")
    b = pac.synth_stress(30000, None, seed=5, lattice_frac=0.1, spacing=2.0,
                         lattice_r=1.70)
    o, _, _ = oracle.run(b, 1.4, 0.0)
    r = pac.compute(b, 1.4, 0.0, chunk_balls=4096)
    _same(r.lists(), o, "stress chunk 4096")
    assert r.stats().sos_calls > 0


def test_chunked_f64_and_canonical():
    _say("case: FP64 12k jittered globule chunk 1024; canonical order")
    b = pac.synth_globule_f64(12000, None, 1.3, seed=4, sigma=0.01)
    one = pac.compute_f64(b, 1.4, 0.0).lists()
    r = pac.compute_f64(b, 1.4, 0.0, chunk_balls=1024)
    _same(r.lists(), one, "f64 chunk 1024")
    c = pac.synth_globule(6000, None, 1.3, seed=3)
    one = pac.compute(c, 1.4, 0.0, order=pac.ORDER_CANONICAL).lists()
    r = pac.compute(c, 1.4, 0.0, order=pac.ORDER_CANONICAL, chunk_balls=700)
    _same(r.lists(), one, "canonical chunk 700")


def test_chunked_single_chunk_and_tiny():
    _say("case: inputs within one chunk; empty input")
    b = pac.synth_globule(1500, None, 1.3, seed=2)
    one = pac.compute(b, 1.4, 0.0).lists()
    _same(pac.compute(b, 1.4, 0.0, chunk_balls=2000).lists(), one, "one chunk")
    r = pac.compute(np.zeros((0, 4), np.float32), 1.4, 0.0, chunk_balls=512)
    assert list(r.counts()) == [0, 0, 0, 0]


def test_chunked_rejects_batches():
    _say("case: chunk_balls on a batch / negative")
    b = pac.synth_globule(600, None, 1.3, seed=2)
    with pytest.raises(pac.PacError) as e:
        pac.compute_batch(b, np.array([0, 300, 600]), 1.4, 0.0, chunk_balls=512)
    assert e.value.code == pac.PAC_ERR_INVALID
    with pytest.raises(pac.PacError) as e:
        pac.compute(b, 1.4, 0.0, chunk_balls=-1)
    assert e.value.code == pac.PAC_ERR_INVALID


def test_chunked_bounds_device_memory():
    """What the LD driver is for: a chunked run's peak device memory is
    that of its largest sub-input, far below one pass over everything."""
    _say("case: 200k globule, peak pool memory one pass vs chunk 16384")
    b = pac.synth_globule(200000, None, 1.3, seed=12)
    pac.memory_trim()
    pac.memory_stats(reset_peak=True)
    one = pac.compute(b, 1.4, 0.0)
    _, peak_one, _ = pac.memory_stats(reset_peak=True)
    lo = one.lists()
    del one
    ch = pac.compute(b, 1.4, 0.0, chunk_balls=16384)
    _, peak_ch, held = pac.memory_stats(reset_peak=True)
    _same(ch.lists(), lo, "200k chunk 16384")
    _say("peak one pass %.1f MB, chunked %.1f MB" % (peak_one / 2**20, peak_ch / 2**20))
    assert 0 < peak_ch < 0.5 * peak_one
    assert held >= peak_ch
    pac.memory_trim()
    u, _, h = pac.memory_stats()
    assert h == u  # nothing cached after a trim


def test_largedata_cli_equals_one_pass(golden_dir, tmp_path):
    crd = os.path.join(golden_dir, "1k4c.crd")
    n = int(open(crd).read().split()[0])
    _say("case: parallelac-largeData 1k4c (%d balls) chunk 3000" % n)
    one = tmp_path / "one.txt"
    big = tmp_path / "big.txt"
    p1 = subprocess.run([os.path.join(BIN, "parallelac"), crd, str(one),
                         "1.4", "0"], capture_output=True, text=True,
                        timeout=120)
    assert p1.returncode == 0, p1.stderr
    p2 = subprocess.run([os.path.join(BIN, "parallelac-largeData"), crd,
                         str(big), "1.4", "0", "3000"], capture_output=True,
                        text=True, timeout=120)
    assert p2.returncode == 0, p2.stderr
    assert big.read_bytes() == one.read_bytes()
    lines = [l for l in p2.stdout.splitlines() if l.startswith("Processed")]
    assert lines == ["Processed %d atoms ... " % (k * 4096)
                     for k in range(_chunks(n, 3000))]
    # the summary counts are the file header's
    head = one.read_text().split("\n", 1)[0].split()
    assert "Number of edges: %18s" % head[1] in p2.stdout
