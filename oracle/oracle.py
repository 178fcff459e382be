"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle
(oracle/_build/libpac_oracle.so, built from oracle/pac_oracle.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product library never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libpac_oracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        L.pac_oracle_run_f32.argtypes = [vp, i64, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(vp)]
        L.pac_oracle_run_f64.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(vp)]
        L.pac_oracle_count.restype = i64
        L.pac_oracle_count.argtypes = [vp, ctypes.c_int]
        L.pac_oracle_copy.argtypes = [vp, ctypes.c_int, vp]
        L.pac_oracle_perm.argtypes = [vp, vp]
        L.pac_oracle_stats.argtypes = [vp, vp]
        L.pac_oracle_times.argtypes = [vp, vp]
        L.pac_oracle_free.argtypes = [vp]
        for f in ("pac_oracle_edge_keep", "pac_oracle_tri_keep",
                  "pac_oracle_tet_keep"):
            getattr(L, f).restype = ctypes.c_int
        L.pac_oracle_edge_keep.argtypes = [vp, vp, ctypes.c_float]
        L.pac_oracle_tri_keep.argtypes = [vp, vp, vp, ctypes.c_float]
        L.pac_oracle_tet_keep.argtypes = [vp, vp, vp, vp, ctypes.c_float]
        L.pac_oracle_tri_violated.argtypes = [vp, vp, vp, vp]
        L.pac_oracle_edge_violated.argtypes = [vp, vp, vp]
        L.pac_oracle_tet_violated.argtypes = [vp, vp, vp]
        L.pac_oracle_sos_det5.argtypes = [vp, vp]
        L.pac_oracle_det5.argtypes = [vp]
        L.pac_oracle_orient3d.argtypes = [vp, vp, vp, vp]
        L.pac_oracle_sos_table.argtypes = [vp, ctypes.c_int]
        _lib = L
    return _lib


STAT_NAMES = ("n", "E_cand", "E_alpha", "T_cand", "T_alpha", "K_cand",
              "K_alpha", "K_final", "T_final", "E_final", "exact_filter",
              "exact_ac2", "sos", "ac2_pairs", "k_enum", "k_ac2", "scale")


def run(xyzr, solvent, alpha, skip_tet_ac2=False, cell_policy=0, nthreads=0):
    """Returns ({1: edges, 2: tris, 3: tets} original ids in reference order,
    stats dict, phase times (ms))."""
    L = lib()
    h = ctypes.c_void_p()
    if np.asarray(xyzr).dtype == np.float64:
        a = np.ascontiguousarray(xyzr, dtype=np.float64).reshape(-1, 4)
        rc = L.pac_oracle_run_f64(a.ctypes.data, a.shape[0], solvent, alpha,
                                  int(skip_tet_ac2), cell_policy, nthreads,
                                  ctypes.byref(h))
    else:
        a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
        rc = L.pac_oracle_run_f32(a.ctypes.data, a.shape[0], solvent, alpha,
                                  int(skip_tet_ac2), cell_policy, nthreads,
                                  ctypes.byref(h))
    if rc != 0:
        raise RuntimeError("oracle failed with status %d" % rc)
    try:
        out = {}
        for d in (1, 2, 3):
            c = L.pac_oracle_count(h, d)
            arr = np.zeros((c, d + 1), dtype=np.int32)
            L.pac_oracle_copy(h, d, arr.ctypes.data)
            out[d] = arr
        st = np.zeros(17, dtype=np.int64)
        L.pac_oracle_stats(h, st.ctypes.data)
        tm = np.zeros(8, dtype=np.float64)
        L.pac_oracle_times(h, tm.ctypes.data)
    finally:
        L.pac_oracle_free(h)
    return out, dict(zip(STAT_NAMES, (int(x) for x in st))), tm


def _f4(b):
    return np.ascontiguousarray(b, dtype=np.float32).reshape(4)


def _ptrs(bs):
    arrs = [_f4(b) for b in bs]
    P = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arrs, P


def edge_keep(b0, b1, alpha):
    a0, a1 = _f4(b0), _f4(b1)
    return lib().pac_oracle_edge_keep(a0.ctypes.data, a1.ctypes.data, alpha)


def tri_keep(b0, b1, b2, alpha):
    a = [_f4(b) for b in (b0, b1, b2)]
    return lib().pac_oracle_tri_keep(*[x.ctypes.data for x in a], alpha)


def tet_keep(b0, b1, b2, b3, alpha):
    a = [_f4(b) for b in (b0, b1, b2, b3)]
    return lib().pac_oracle_tet_keep(*[x.ctypes.data for x in a], alpha)


def tri_violated(b0, b1, b2, e):
    a = [_f4(b) for b in (b0, b1, b2, e)]
    return lib().pac_oracle_tri_violated(*[x.ctypes.data for x in a])


def edge_violated(b0, b1, e):
    a = [_f4(b) for b in (b0, b1, e)]
    return lib().pac_oracle_edge_violated(*[x.ctypes.data for x in a])


def tet_violated(b4, e, ids5):
    arrs, P = _ptrs(b4)
    ee = _f4(e)
    ids = np.ascontiguousarray(ids5, dtype=np.int64)
    return lib().pac_oracle_tet_violated(P, ee.ctypes.data, ids.ctypes.data)


def sos_det5(b5, ids5):
    arrs, P = _ptrs(b5)
    ids = np.ascontiguousarray(ids5, dtype=np.int64)
    return lib().pac_oracle_sos_det5(P, ids.ctypes.data)


def det5(b5):
    arrs, P = _ptrs(b5)
    return lib().pac_oracle_det5(P)


def orient3d(a, b, c, d):
    x = [_f4(v) for v in (a, b, c, d)]
    return lib().pac_oracle_orient3d(*[v.ctypes.data for v in x])


def sos_table():
    out = np.zeros((64, 3), dtype=np.int32)
    n = lib().pac_oracle_sos_table(out.ctypes.data, 64)
    return out[:n]
