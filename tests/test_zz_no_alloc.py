"""GPU: the allocation-free entry point pac_compute_noalloc (device scratch
array, caller-provided lists) through the C ABI, against pac_compute on the
same inputs: counts and lists identical, row order included, on two calls
in a row (the second with a smaller input, so stale scratch contents from
the first would show), with device and host buffers, both row orders, and
its error paths (too many balls, lists over capacity).
"""
import ctypes
import sys

import numpy as np
import pytest

import pac

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(xyzr_ptr, n, solvent, alpha, outs, caps, **kw):
    """call pac_compute_noalloc; returns (status, counts)"""
    o = pac.options(**kw)
    counts = np.zeros(4, dtype=np.int64)
    rc = pac.lib().pac_compute_noalloc(
        ctypes.c_void_p(xyzr_ptr), n, solvent, alpha, ctypes.byref(o),
        ctypes.c_void_p(outs[0]), caps[0], ctypes.c_void_p(outs[1]), caps[1],
        ctypes.c_void_p(outs[2]), caps[2], counts.ctypes.data)
    return rc, counts


def _reference(b, solvent, alpha, **kw):
    r = pac.compute(b, solvent, alpha, **kw)
    return r.counts(), r.lists()


def _device_run(torch, b, solvent, alpha, slack, **kw):
    """device input, device lists with `slack` spare rows each"""
    want_c, want = _reference(b, solvent, alpha, **kw)
    dev = torch.from_numpy(np.ascontiguousarray(b)).to("cuda:0")
    outs = [torch.full((int(want_c[d]) + slack, d + 1), -7, dtype=torch.int32,
                       device="cuda:0") for d in (1, 2, 3)]
    torch.cuda.synchronize()
    rc, counts = _noalloc(dev.data_ptr(), b.shape[0], solvent, alpha,
                          [t.data_ptr() for t in outs],
                          [t.shape[0] for t in outs], device_inputs=True, **kw)
    assert rc == pac.PAC_OK, pac.lib().pac_last_error()
    assert counts.tolist() == want_c.tolist()
    for d in (1, 2, 3):
        got = outs[d - 1].cpu().numpy()
        m = int(counts[d])
        assert np.array_equal(got[:m], want[d]), "dim %d differs" % d
        assert (got[m:] == -7).all(), "dim %d: rows past the count written" % d
    return counts


def test_noalloc_two_calls_device_buffers():
    torch = pytest.importorskip("torch")
    big = pac.synth_globule(2500, None, 1.3, seed=21)
    small = pac.synth_globule(600, None, 1.3, seed=22)
    _say("noalloc: 2500 balls, device buffers")
    c1 = _device_run(torch, big, 1.4, 0.0, 5)
    _say("noalloc: 600 balls after it")
    c2 = _device_run(torch, small, 1.4, 0.0, 3)
    assert c2[3] < c1[3]


def test_noalloc_host_buffers_and_options():
    torch = pytest.importorskip("torch")  # noqa: F841  (HIP runtime from torch)
    b = pac.synth_globule(1500, None, 1.3, seed=23)
    for kw in ({"order": pac.ORDER_CANONICAL},
               {"precision": pac.PREC_PLAIN_FP64, "cell_policy": pac.CELL_REFERENCE_RAW},
               {}):
        _say("noalloc: host buffers %s" % kw)
        want_c, want = _reference(b, 1.4, 0.5, **kw)
        outs = [np.zeros((int(want_c[d]), d + 1), dtype=np.int32) for d in (1, 2, 3)]
        rc, counts = _noalloc(b.ctypes.data, b.shape[0], 1.4, 0.5,
                              [a.ctypes.data for a in outs],
                              [a.shape[0] for a in outs], **kw)
        assert rc == pac.PAC_OK, pac.lib().pac_last_error()
        assert counts.tolist() == want_c.tolist()
        for d in (1, 2, 3):
            assert np.array_equal(outs[d - 1], want[d]), "dim %d differs" % d


def test_noalloc_errors():
    torch = pytest.importorskip("torch")  # noqa: F841
    b = pac.synth_globule(800, None, 1.3, seed=24)
    want_c, _ = _reference(b, 1.4, 0.0)
    # lists over capacity: nothing written, the sizes needed reported
    outs = [np.full((4, d + 1), -3, dtype=np.int32) for d in (1, 2, 3)]
    rc, counts = _noalloc(b.ctypes.data, b.shape[0], 1.4, 0.0,
                          [a.ctypes.data for a in outs], [4, 4, 4])
    assert rc == pac.PAC_ERR_OVERFLOW
    assert counts.tolist() == want_c.tolist()
    assert all((a == -3).all() for a in outs)
    # too many balls: refused before any device work
    n = pac.NOALLOC_MAX_BALLS + 1
    big = np.zeros((n, 4), dtype=np.float32)
    rc, _ = _noalloc(big.ctypes.data, n, 1.4, 0.0, [0, 0, 0], [0, 0, 0])
    assert rc == pac.PAC_ERR_INVALID
    assert b"maximum" in pac.lib().pac_last_error()
    # options that need allocation are refused
    rc, _ = _noalloc(b.ctypes.data, b.shape[0], 1.4, 0.0, [0, 0, 0], [0, 0, 0],
                     chunk_balls=512)
    assert rc == pac.PAC_ERR_INVALID
    # an empty input
    rc, counts = _noalloc(b.ctypes.data, 0, 1.4, 0.0, [0, 0, 0], [0, 0, 0])
    assert rc == pac.PAC_OK and counts.tolist() == [0, 0, 0, 0]
    # after the errors, a normal call still matches
    outs = [np.zeros((int(want_c[d]), d + 1), dtype=np.int32) for d in (1, 2, 3)]
    rc, counts = _noalloc(b.ctypes.data, b.shape[0], 1.4, 0.0,
                          [a.ctypes.data for a in outs],
                          [a.shape[0] for a in outs])
    assert rc == pac.PAC_OK and counts.tolist() == want_c.tolist()
