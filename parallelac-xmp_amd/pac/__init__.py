"""Python binding of libpac (the MI355X weighted alpha complex engine).

Thin ctypes layer over the C ABI in include/pac.h.  The reference
(forest113/ParallelAC-xmp) exposes only the `parallelac` CLI
(AlphaComplexParallel.cu:1053-1057); `alpha_complex()` is the in-process
equivalent: balls in (x, y, z, r_raw), solvent and alpha, sorted edge /
triangle / tetrahedron lists out in the reference's row order.

There is no CPU fallback: if libpac.so is missing or no GPU is visible the
call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# the in-tree library, unless PAC_LIB names another build of it
LIB_PATH = os.environ.get("PAC_LIB") or os.path.join(os.path.dirname(_HERE),
                                                      "lib", "libpac.so")

PAC_OK = 0
PAC_ERR_IO, PAC_ERR_INVALID, PAC_ERR_DEGENERATE = -1, -2, -3
PAC_ERR_NOMEM, PAC_ERR_DEVICE, PAC_ERR_OVERFLOW, PAC_ERR_RANGE = -4, -5, -6, -7
PAC_ERR_STALE = -8
CELL_EXACT, CELL_REFERENCE_RAW = 0, 1
PREC_PLAIN_FP64, PREC_STAGED = 0, 1
ORDER_REFERENCE, ORDER_CANONICAL = 0, 1

PHASE_LABELS = ("Data transfer", "Grid computation", "Potential edges",
                "Potential triangles", "Potential tetrahedra",
                "AC2 test for tetrahedra", "AC2 test for triangles",
                "AC2 test for edges")
FAMILIES = ("edge_filter", "tri_filter", "tet_filter", "tet_ac2", "tri_ac2",
            "edge_ac2")
STAGES = ("fp32", "fp64", "dd", "wide_int")
KERNELS = ("edge_fill", "tri_fill", "tet_fill", "tet_ac2", "tri_ac2",
           "edge_ac2", "radix_sort", "tet_ac2_exact")


class Options(ctypes.Structure):
    _fields_ = [("cell_policy", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("skip_tet_ac2", ctypes.c_int32), ("order", ctypes.c_int32),
                ("device_inputs", ctypes.c_int32), ("device", ctypes.c_int32),
                ("stream", ctypes.c_void_p), ("block", ctypes.c_int32),
                ("profile", ctypes.c_int32), ("ac2_budget", ctypes.c_int32),
                ("batch_balls", ctypes.c_int32), ("ac2_grid", ctypes.c_int32),
                ("ac2_fine", ctypes.c_int32), ("ac2_quick", ctypes.c_int32),
                ("chunk_balls", ctypes.c_int32), ("filtration", ctypes.c_int32),
                ("ac2_chunk", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64),
                ("edge_candidates", ctypes.c_int64), ("edge_alpha", ctypes.c_int64),
                ("edges_final", ctypes.c_int64),
                ("tri_candidates", ctypes.c_int64), ("tri_alpha", ctypes.c_int64),
                ("tris_final", ctypes.c_int64),
                ("tet_candidates", ctypes.c_int64), ("tet_alpha", ctypes.c_int64),
                ("tets_final", ctypes.c_int64),
                ("stage_evals", (ctypes.c_int64 * 4) * 6),
                ("sos_calls", ctypes.c_int64),
                ("ac2_pairs", ctypes.c_int64 * 3),
                ("k_enum", ctypes.c_int32), ("k_ac2", ctypes.c_int32),
                ("scale_bits", ctypes.c_int32), ("grid_cells", ctypes.c_int32),
                ("phase_ms", ctypes.c_double * 8), ("total_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double * 8)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_
             if f not in ("stage_evals", "phase_ms", "ac2_pairs", "kernel_ms")}
        d["stage_evals"] = {FAMILIES[f]: {STAGES[s]: int(self.stage_evals[f][s])
                                          for s in range(4)} for f in range(6)}
        d["phase_ms"] = {PHASE_LABELS[i]: float(self.phase_ms[i]) for i in range(8)}
        d["ac2_pairs"] = [int(x) for x in self.ac2_pairs]
        d["kernel_ms"] = {KERNELS[i]: float(self.kernel_ms[i]) for i in range(8)}
        return d


# every symbol include/pac.h declares
EXPORTS = (
    "pac_options_init", "pac_compute", "pac_compute_f64", "pac_result_counts",
    "pac_result_copy_edges", "pac_result_copy_triangles", "pac_result_copy_tets",
    "pac_result_device_list", "pac_result_copy_alpha", "pac_result_stats",
    "pac_result_free", "pac_compute_batch", "pac_batch_counts", "pac_batch_copy",
    "pac_batch_stats", "pac_batch_free", "pac_alpha_adjoint", "pac_read_crd",
    "pac_write_output", "pac_free", "pac_synth_globule", "pac_synth_globule_f64",
    "pac_synth_batch", "pac_synth_stress", "pac_last_error", "pac_version",
    "pac_autotune", "pac_graph_create", "pac_graph_set_input", "pac_graph_launch",
    "pac_graph_result", "pac_graph_info", "pac_graph_free", "pac_write_crd",
    "pac_read_pdb", "pac_result_copy_birth", "pac_result_copy_filtration",
    "pac_memory_stats", "pac_memory_trim", "pac_compute_noalloc",
)
# the allocation-free entry point's limits (include/pac.h)
NOALLOC_SCRATCH_BYTES = 256 << 20
NOALLOC_MAX_BALLS = 8192

_lib = None


def lib():
    """Load libpac.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpac.so not built: run __graft_entry__.build() "
                           "(or make -C parallelac-xmp_amd)")
    # PyTorch-ROCm bundles its own HIP runtime under the same soname as
    # ROCm's.  Loading torch first makes libpac bind to that copy, so the
    # process holds one HIP runtime; loaded the other way round, torch's
    # second copy sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    L.pac_options_init.argtypes = [ctypes.POINTER(Options)]
    L.pac_compute.argtypes = [vp, i64, ctypes.c_float, ctypes.c_float,
                              ctypes.POINTER(Options), ctypes.POINTER(vp)]
    L.pac_compute_f64.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double,
                                  ctypes.POINTER(Options), ctypes.POINTER(vp)]
    L.pac_compute_noalloc.argtypes = [vp, i64, ctypes.c_float, ctypes.c_float,
                                      ctypes.POINTER(Options), vp, i64, vp, i64,
                                      vp, i64, vp]
    L.pac_result_counts.argtypes = [vp, vp]
    for f in ("pac_result_copy_edges", "pac_result_copy_triangles",
              "pac_result_copy_tets"):
        getattr(L, f).argtypes = [vp, vp]
    L.pac_result_device_list.argtypes = [vp, ctypes.c_int]
    L.pac_result_device_list.restype = vp
    L.pac_result_copy_alpha.argtypes = [vp, ctypes.c_int, vp]
    L.pac_result_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.pac_result_copy_birth.argtypes = [vp, ctypes.c_int, vp]
    L.pac_memory_stats.argtypes = [ctypes.c_int, ctypes.POINTER(i64),
                                   ctypes.POINTER(i64), ctypes.POINTER(i64),
                                   ctypes.c_int]
    L.pac_memory_trim.argtypes = [ctypes.c_int]
    L.pac_result_copy_filtration.argtypes = [vp, vp, vp, vp]
    L.pac_result_free.argtypes = [vp]
    L.pac_compute_batch.argtypes = [vp, vp, i64, ctypes.c_float, ctypes.c_float,
                                    ctypes.POINTER(Options), ctypes.POINTER(vp)]
    L.pac_batch_counts.argtypes = [vp, vp]
    L.pac_batch_copy.argtypes = [vp, i64, ctypes.c_int, vp]
    L.pac_batch_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.pac_batch_free.argtypes = [vp]
    L.pac_alpha_adjoint.argtypes = [vp, vp, vp, vp, vp]
    L.pac_autotune.argtypes = [vp, i64, vp, i64, ctypes.c_float, ctypes.c_float,
                               ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_double)]
    L.pac_graph_create.argtypes = [vp, vp, i64, ctypes.c_float, ctypes.c_float,
                                   ctypes.POINTER(Options), ctypes.c_double,
                                   ctypes.POINTER(vp)]
    L.pac_graph_set_input.argtypes = [vp, vp, vp]
    L.pac_graph_launch.argtypes = [vp, vp]
    L.pac_graph_result.argtypes = [vp, ctypes.POINTER(vp)]
    L.pac_graph_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.pac_graph_free.argtypes = [vp]
    L.pac_read_crd.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp),
                               ctypes.POINTER(i64)]
    L.pac_write_output.argtypes = [ctypes.c_char_p, vp, i64, ctypes.c_float, vp]
    L.pac_free.argtypes = [vp]
    L.pac_write_crd.argtypes = [ctypes.c_char_p, vp, i64]
    L.pac_read_pdb.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                               ctypes.POINTER(vp), ctypes.POINTER(i64)]
    L.pac_synth_globule.argtypes = [i64, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_uint64, ctypes.c_int, vp]
    L.pac_synth_globule_f64.argtypes = [i64, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, ctypes.c_double, vp]
    L.pac_synth_batch.argtypes = [i64, i64, i64, ctypes.c_uint64, vp, vp]
    L.pac_synth_stress.argtypes = [i64, ctypes.c_double, ctypes.c_uint64,
                                   ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, vp]
    L.pac_last_error.restype = ctypes.c_char_p
    L.pac_version.restype = ctypes.c_char_p
    _lib = L
    return L


class PacError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("pac error %d: %s" % (code, msg))
        self.code = code


def _check(rc):
    if rc != PAC_OK:
        raise PacError(rc, lib().pac_last_error().decode(errors="replace"))


def options(cell_policy=CELL_EXACT, precision=PREC_STAGED, skip_tet_ac2=False,
            order=ORDER_REFERENCE, device_inputs=False, device=0, stream=None,
            profile=False, ac2_budget=0, batch_balls=0, ac2_grid=0, ac2_fine=0,
            ac2_quick=0, chunk_balls=0, filtration=False, block=0, ac2_chunk=0):
    o = Options()
    lib().pac_options_init(ctypes.byref(o))
    o.cell_policy = cell_policy
    o.precision = precision
    o.skip_tet_ac2 = int(bool(skip_tet_ac2))
    o.order = order
    o.device_inputs = int(bool(device_inputs))
    o.device = device
    o.stream = stream
    o.profile = int(bool(profile))
    o.ac2_budget = int(ac2_budget)
    o.batch_balls = int(batch_balls)
    o.ac2_grid = int(ac2_grid)
    o.ac2_fine = int(ac2_fine)
    o.ac2_quick = int(ac2_quick)
    o.chunk_balls = int(chunk_balls)
    o.filtration = int(bool(filtration))
    o.block = int(block)
    o.ac2_chunk = int(ac2_chunk)
    return o


class Result:
    """Owned handle to a pac_result; lists are copied out on demand."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.pac_result_free(self._h)
            self._h = ctypes.c_void_p(None)

    @property
    def handle(self):
        return self._h

    def counts(self):
        c = np.zeros(4, dtype=np.int64)
        _check(lib().pac_result_counts(self._h, c.ctypes.data))
        return c

    def lists(self):
        c = self.counts()
        out = {}
        fns = (lib().pac_result_copy_edges, lib().pac_result_copy_triangles,
               lib().pac_result_copy_tets)
        for d in (1, 2, 3):
            a = np.zeros((int(c[d]), d + 1), dtype=np.int32)
            _check(fns[d - 1](self._h, a.ctypes.data))
            out[d] = a
        return out

    def device_list(self, dim):
        return lib().pac_result_device_list(self._h, dim)

    def stats(self):
        s = Stats()
        _check(lib().pac_result_stats(self._h, ctypes.byref(s)))
        return s

    def adjoint(self, g_tri=None, g_tet=None):
        """dL/d(x,y,z) (n x 3) and dL/dr (n) from dL/drho of the output
        triangles and tets (FP64, list order; None = zero)."""
        c = self.counts()
        n = int(c[0])
        args = []
        for g, m in ((g_tri, c[2]), (g_tet, c[3])):
            if g is None:
                args.append(None)
            else:
                a = np.ascontiguousarray(g, dtype=np.float64)
                assert a.shape == (int(m),)
                args.append(a)
        dxyz = np.zeros((n, 3), dtype=np.float64)
        dr = np.zeros(n, dtype=np.float64)
        _check(lib().pac_alpha_adjoint(
            self._h, None if args[0] is None else args[0].ctypes.data,
            None if args[1] is None else args[1].ctypes.data,
            dxyz.ctypes.data, dr.ctypes.data))
        return dxyz, dr

    def alpha_values(self, dim):
        n = int(self.counts()[dim])
        rho = np.zeros(n, dtype=np.float64)
        _check(lib().pac_result_copy_alpha(self._h, dim, rho.ctypes.data))
        return rho

    def copy_list_to(self, dim, ptr):
        """copy the dim-1/2/3 list into device or host memory at ptr
        (counts()[dim] x (dim+1) int32)"""
        fns = (lib().pac_result_copy_edges, lib().pac_result_copy_triangles,
               lib().pac_result_copy_tets)
        if int(self.counts()[dim]):
            _check(fns[dim - 1](self._h, ctypes.c_void_p(ptr)))

    def copy_alpha_to(self, dim, ptr):
        """rho of the dim-1/2/3 simplices into device or host memory at ptr
        (counts()[dim] float64)"""
        if int(self.counts()[dim]):
            _check(lib().pac_result_copy_alpha(self._h, dim, ctypes.c_void_p(ptr)))

    def adjoint_to(self, g_tri_ptr, g_tet_ptr, dxyz_ptr, dr_ptr):
        """pac_alpha_adjoint on raw (host or device) pointers; a gradient
        pointer may be None (zero)"""
        vp = ctypes.c_void_p
        _check(lib().pac_alpha_adjoint(self._h, vp(g_tri_ptr), vp(g_tet_ptr),
                                       vp(dxyz_ptr), vp(dr_ptr)))

    def births(self, dim):
        """Filtration values of the simplices of dim 0..3 (needs
        filtration=True at compute time)."""
        n = int(self.counts()[dim])
        v = np.zeros(n, dtype=np.float64)
        _check(lib().pac_result_copy_birth(self._h, dim, v.ctypes.data))
        return v

    def filtration(self):
        """All simplices sorted by filtration value: (dim, index, value)."""
        m = int(sum(self.counts()))
        d = np.zeros(m, dtype=np.int32)
        i = np.zeros(m, dtype=np.int64)
        v = np.zeros(m, dtype=np.float64)
        _check(lib().pac_result_copy_filtration(self._h, d.ctypes.data,
                                                i.ctypes.data, v.ctypes.data))
        return d, i, v


def compute(xyzr, solvent, alpha, **kw):
    """Run the pipeline on host FP32 balls (n x 4: x, y, z, r_raw)."""
    a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
    h = ctypes.c_void_p()
    o = options(**kw)
    _check(lib().pac_compute(a.ctypes.data, a.shape[0], solvent, alpha,
                             ctypes.byref(o), ctypes.byref(h)))
    return Result(h.value)


def compute_f64(xyzr, solvent, alpha, **kw):
    """Run the pipeline on host FP64 balls (continuous coordinates)."""
    a = np.ascontiguousarray(xyzr, dtype=np.float64).reshape(-1, 4)
    h = ctypes.c_void_p()
    o = options(**kw)
    _check(lib().pac_compute_f64(a.ctypes.data, a.shape[0], solvent, alpha,
                                 ctypes.byref(o), ctypes.byref(h)))
    return Result(h.value)


def compute_device(ptr, n, solvent, alpha, **kw):
    """Run on a device-resident FP32 n x 4 array (e.g. a torch tensor)."""
    h = ctypes.c_void_p()
    o = options(device_inputs=True, **kw)
    _check(lib().pac_compute(ctypes.c_void_p(ptr), n, solvent, alpha,
                             ctypes.byref(o), ctypes.byref(h)))
    return Result(h.value)


def compute_device_f64(ptr, n, solvent, alpha, **kw):
    """Run on a device-resident FP64 n x 4 array."""
    h = ctypes.c_void_p()
    o = options(device_inputs=True, **kw)
    _check(lib().pac_compute_f64(ctypes.c_void_p(ptr), n, solvent, alpha,
                                 ctypes.byref(o), ctypes.byref(h)))
    return Result(h.value)


class BatchResult:
    """Owned handle to a pac_batch_result (per-molecule lists, local ids)."""

    def __init__(self, handle, n_mol):
        self._h = ctypes.c_void_p(handle)
        self.n_mol = n_mol

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.pac_batch_free(self._h)
            self._h = ctypes.c_void_p(None)

    def counts(self):
        c = np.zeros((self.n_mol, 4), dtype=np.int64)
        _check(lib().pac_batch_counts(self._h, c.ctypes.data))
        return c

    def lists(self, m):
        c = self.counts()[m]
        out = {}
        for d in (1, 2, 3):
            a = np.zeros((int(c[d]), d + 1), dtype=np.int32)
            _check(lib().pac_batch_copy(self._h, m, d, a.ctypes.data))
            out[d] = a
        return out

    def stats(self):
        s = Stats()
        _check(lib().pac_batch_stats(self._h, ctypes.byref(s)))
        return s


def compute_batch(xyzr, offsets, solvent, alpha, **kw):
    """Batched ensemble: xyzr (host n x 4 FP32, or a device pointer with
    device_inputs=True and n given by offsets[-1]) and int64 offsets."""
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n_mol = off.shape[0] - 1
    h = ctypes.c_void_p()
    o = options(**kw)
    if kw.get("device_inputs"):
        ptr = ctypes.c_void_p(int(xyzr))
    else:
        a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
        ptr = a.ctypes.data
    _check(lib().pac_compute_batch(ptr, off.ctypes.data, n_mol, solvent, alpha,
                                   ctypes.byref(o), ctypes.byref(h)))
    return BatchResult(h.value, n_mol)


class Graph:
    """A batch recorded as one HIP graph (pac_graph_*): launch() runs the
    whole batch with one hipGraphLaunch, result() returns a BatchResult."""

    def __init__(self, xyzr, offsets, solvent, alpha, headroom=-1.0, **kw):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.n_mol = off.shape[0] - 1
        self._n = int(off[-1])
        self._h = ctypes.c_void_p()
        o = options(**kw)
        if kw.get("device_inputs"):
            ptr = ctypes.c_void_p(int(xyzr))
        else:
            a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
            ptr = a.ctypes.data
        _check(lib().pac_graph_create(ptr, off.ctypes.data, self.n_mol, solvent,
                                      alpha, ctypes.byref(o), headroom,
                                      ctypes.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.pac_graph_free(self._h)
            self._h = ctypes.c_void_p(None)

    def set_input(self, xyzr, stream=None):
        """new coordinates (host n x 4 FP32 array, or an int device pointer)"""
        if isinstance(xyzr, int):
            ptr = ctypes.c_void_p(xyzr)
        else:
            a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
            if a.shape[0] != self._n:
                raise ValueError("ball count differs from the recorded batch")
            ptr = a.ctypes.data
        _check(lib().pac_graph_set_input(self._h, ptr, stream))

    def launch(self, stream=None):
        _check(lib().pac_graph_launch(self._h, stream))

    def result(self):
        h = ctypes.c_void_p()
        _check(lib().pac_graph_result(self._h, ctypes.byref(h)))
        return BatchResult(h.value, self.n_mol)

    def info(self):
        """(kernel and other nodes in the graph, device bytes it holds)"""
        nodes, nbytes = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().pac_graph_info(self._h, ctypes.byref(nodes), ctypes.byref(nbytes)))
        return nodes.value, nbytes.value


def autotune(xyzr, solvent, alpha, offsets=None, **kw):
    """Pick the fastest ac2_budget / ac2_grid for these balls (host FP32
    n x 4, or a device pointer with device_inputs=True and n given as
    kw['n']); returns (options dict for compute(), best ms)."""
    n = kw.pop("n", None)
    o = options(**kw)
    if kw.get("device_inputs"):
        ptr = ctypes.c_void_p(int(xyzr))
    else:
        a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
        ptr, n = a.ctypes.data, a.shape[0]
    best = ctypes.c_double()
    if offsets is None:
        _check(lib().pac_autotune(ptr, n, None, 0, solvent, alpha, ctypes.byref(o),
                                  ctypes.byref(best)))
    else:
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        _check(lib().pac_autotune(ptr, int(off[-1]), off.ctypes.data, off.shape[0] - 1,
                                  solvent, alpha, ctypes.byref(o), ctypes.byref(best)))
    return ({"ac2_budget": o.ac2_budget, "ac2_grid": o.ac2_grid,
             "ac2_fine": o.ac2_fine, "ac2_quick": o.ac2_quick,
             "block": o.block, "ac2_chunk": o.ac2_chunk}, best.value)


def alpha_complex(xyzr, solvent=0.0, alpha=0.0, **kw):
    """Convenience: returns (edges, triangles, tets) int32 arrays."""
    r = compute(xyzr, solvent, alpha, **kw)
    L = r.lists()
    return L[1], L[2], L[3]


def _take_balls(p, n):
    try:
        buf = (ctypes.c_float * (4 * n.value)).from_address(p.value) if n.value else []
        a = np.array(buf, dtype=np.float32).reshape(-1, 4)
    finally:
        lib().pac_free(p)
    return a


def memory_stats(device=0, reset_peak=False):
    """Device bytes of the library's pool: (in_use, peak, held)."""
    u, p, h = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _check(lib().pac_memory_stats(device, ctypes.byref(u), ctypes.byref(p),
                                  ctypes.byref(h), int(bool(reset_peak))))
    return u.value, p.value, h.value


def memory_trim(device=0):
    _check(lib().pac_memory_trim(device))


def read_crd(path):
    p = ctypes.c_void_p()
    n = ctypes.c_int64()
    _check(lib().pac_read_crd(path.encode(), ctypes.byref(p), ctypes.byref(n)))
    return _take_balls(p, n)


PDB_HETATM, PDB_HYDROGENS, PDB_WATER = 1, 2, 4


def read_pdb(path, hetatm=False, hydrogens=False, water=False, radii=None):
    """Balls (x, y, z, r) of a PDB file (pac_read_pdb).  radii: dict or
    "E=r,..." string overriding the element radii ("*": any other)."""
    flags = ((PDB_HETATM if hetatm else 0) | (PDB_HYDROGENS if hydrogens else 0)
             | (PDB_WATER if water else 0))
    if isinstance(radii, dict):
        radii = ",".join("%s=%r" % (k, float(v)) for k, v in radii.items())
    p = ctypes.c_void_p()
    n = ctypes.c_int64()
    _check(lib().pac_read_pdb(path.encode(), flags,
                              radii.encode() if radii else None,
                              ctypes.byref(p), ctypes.byref(n)))
    return _take_balls(p, n)


def write_crd(path, xyzr):
    """CRD file in the layout of the reference's data files."""
    a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
    _check(lib().pac_write_crd(path.encode(), a.ctypes.data, a.shape[0]))


def write_output(path, xyzr, solvent, result):
    a = np.ascontiguousarray(xyzr, dtype=np.float32).reshape(-1, 4)
    _check(lib().pac_write_output(path.encode(), a.ctypes.data, a.shape[0],
                                  solvent, result.handle))


# ---------------------------------------------------------------- synth
def synth_globule(n, sphere_radius=None, min_sep=1.3, seed=1, decimals=3):
    if sphere_radius is None:
        sphere_radius = (3.0 * n / (4.0 * np.pi * 0.05)) ** (1.0 / 3.0)
    out = np.zeros((n, 4), dtype=np.float32)
    _check(lib().pac_synth_globule(n, sphere_radius, min_sep, seed, decimals,
                                   out.ctypes.data))
    return out


def synth_globule_f64(n, sphere_radius=None, min_sep=1.3, seed=4, sigma=0.01):
    if sphere_radius is None:
        sphere_radius = (3.0 * n / (4.0 * np.pi * 0.05)) ** (1.0 / 3.0)
    out = np.zeros((n, 4), dtype=np.float64)
    _check(lib().pac_synth_globule_f64(n, sphere_radius, min_sep, seed, sigma,
                                       out.ctypes.data))
    return out


def synth_batch(n_mol, nmin=200, nmax=20000, seed=3):
    off = np.zeros(n_mol + 1, dtype=np.int64)
    _check(lib().pac_synth_batch(n_mol, nmin, nmax, seed, off.ctypes.data, None))
    out = np.zeros((int(off[-1]), 4), dtype=np.float32)
    _check(lib().pac_synth_batch(n_mol, nmin, nmax, seed, off.ctypes.data,
                                 out.ctypes.data))
    return out, off


def synth_stress(n, sphere_radius=None, seed=5, lattice_frac=0.1, spacing=2.0,
                 lattice_r=1.70):
    if sphere_radius is None:
        sphere_radius = (3.0 * n / (4.0 * np.pi * 0.05)) ** (1.0 / 3.0)
    out = np.zeros((n, 4), dtype=np.float32)
    _check(lib().pac_synth_stress(n, sphere_radius, seed, lattice_frac, spacing,
                                  lattice_r, out.ctypes.data))
    return out
