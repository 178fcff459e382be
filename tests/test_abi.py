"""CPU: the C-ABI library loads, exports everything include/pac.h declares,
and its host-side parts (CRD reader, generators) behave; no GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import pac

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "pac.h")).read()
    declared = set(re.findall(r"\b(pac_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(pac.EXPORTS)
    L = pac.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert L.pac_version().startswith(b"parallelac")


def test_read_crd_fixture(golden_dir):
    b = pac.read_crd(os.path.join(golden_dir, "1grm.crd"))
    assert b.shape == (260, 4)
    assert b.dtype == np.float32
    assert b[0].tolist() == [np.float32(2.1946), np.float32(-2.7738),
                             np.float32(-0.4134), np.float32(1.4)]
    b2 = pac.read_crd(os.path.join(golden_dir, "1k4c.crd"))  # tabs + '#TER'
    assert b2.shape == (16068, 4)


def test_read_crd_errors(tmp_path):
    with pytest.raises(pac.PacError) as e:
        pac.read_crd(str(tmp_path / "missing.crd"))
    assert e.value.code == pac.PAC_ERR_IO
    p = tmp_path / "short.crd"
    p.write_text("3\n1 2 3 1.5\n4 5 6\n")
    with pytest.raises(pac.PacError) as e:
        pac.read_crd(str(p))
    assert e.value.code == pac.PAC_ERR_IO
    p2 = tmp_path / "empty.crd"
    p2.write_text("0\n")
    assert pac.read_crd(str(p2)).shape == (0, 4)


def test_synth_globule_properties():
    a = pac.synth_globule(3000, None, 1.3, seed=1)
    b = pac.synth_globule(3000, None, 1.3, seed=1)
    assert np.array_equal(a, b)
    R = (3 * 3000 / (4 * np.pi * 0.05)) ** (1 / 3)
    r = np.linalg.norm(a[:, :3].astype(np.float64), axis=1)
    assert r.max() <= R + 0.001
    # coordinates rounded to 0.001
    assert np.allclose(np.round(a[:, :3].astype(np.float64) * 1000), a[:, :3] * 1000, atol=1e-3)
    assert set(np.unique(a[:, 3]).tolist()) <= {np.float32(v) for v in (1.70, 1.55, 1.52, 1.80)}
    from scipy.spatial import cKDTree
    d, _ = cKDTree(a[:, :3]).query(a[:, :3], k=2)
    assert d[:, 1].min() >= 1.3 - 0.002


def test_synth_batch_and_stress():
    x, off = pac.synth_batch(8, 200, 2000, seed=3)
    assert off[0] == 0 and len(off) == 9 and off[-1] == len(x)
    assert ((np.diff(off) >= 200) & (np.diff(off) <= 2000)).all()
    s = pac.synth_stress(20000, None, seed=5)
    assert s.shape == (20000, 4)
    assert (s[:, 3] == np.float32(1.70)).sum() >= 1728


def test_batch_argument_errors():
    """offsets are validated on the host before any device work"""
    xyzr = np.zeros((4, 4), dtype=np.float32)
    for bad in ([0, 3, 2, 4], [1, 4]):
        with pytest.raises(pac.PacError) as e:
            pac.compute_batch(xyzr, np.array(bad), 1.4, 0.0)
        assert e.value.code == pac.PAC_ERR_INVALID


def test_graph_argument_errors():
    """recorded batches validate on the host before any device work"""
    xyzr = np.zeros((4, 4), dtype=np.float32)
    off = np.array([0, 4])
    for kw in (dict(order=pac.ORDER_CANONICAL), dict(headroom=100.0)):
        with pytest.raises(pac.PacError) as e:
            pac.Graph(xyzr, off, 1.4, 0.0, **kw)
        assert e.value.code == pac.PAC_ERR_INVALID
    with pytest.raises(pac.PacError) as e:
        pac.Graph(xyzr, np.array([0, 3, 2, 4]), 1.4, 0.0)
    assert e.value.code == pac.PAC_ERR_INVALID


def test_memory_stats_without_a_device():
    """A device the library never used (no engine): the pool reads zero
    and trim is a no-op."""
    dev = 7
    assert pac.memory_stats(dev) == (0, 0, 0)
    assert pac.memory_stats(dev, reset_peak=True) == (0, 0, 0)
    pac.memory_trim(dev)
    u, p, h = pac.memory_stats(0)
    assert 0 <= u <= p and u <= h
