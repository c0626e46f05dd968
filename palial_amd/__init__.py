"""palial_amd -- MI355X (gfx950) tiled dense factorizations.

Python mirror of the reference's algorithm interface over the C ABI of
``libpalial_hip.so`` (include/palial_hip.h):

    cholesky(A)   <- void cholesky(MATRIX_desc A)            src/cholesky.h:12
    lu(A)         <- void lu(MATRIX_desc A)                  src/lu.h:11
    qr(A, S)      <- void qr(MATRIX_desc A, MATRIX_desc S)   src/qr.h:15

``A`` may be a MATRIX_desc, a flat float64 numpy array in tile layout L (host
memory, factored in place), or a float64 CUDA/HIP torch tensor (device memory,
factored in place).  Tile kernels (dpotrf, dgetrf, dgeqrt, dormqr, dtsqrt,
dtsmqr and the BLAS tile tasks) take column-major n x n numpy tiles.

There is no CPU fallback: every call goes to the HIP library, and importing
the package fails loudly when ``libpalial_hip.so`` has not been built.
"""
import atexit
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpalial_hip.so")
DRIVER_PATH = os.path.join(_HERE, "bin", "palial")

PREC_FP64, PREC_FP32, PREC_BF16, PREC_FP8, PREC_BF16X3 = 0, 1, 2, 3, 4
LEFT, RIGHT = 141, 142
NOTRANS, TRANS = 111, 112
TASK_CHOL_TRSM, TASK_CHOL_SYRK, TASK_CHOL_GEMM = 1, 2, 3
TASK_LU_TRSM_R, TASK_LU_TRSM_L, TASK_LU_GEMM = 4, 5, 6
GEN_SYM, GEN_DENSE, GEN_PALIAL, GEN_IDENTITY, GEN_GAUSS, GEN_GEOM_DIAG = 0, 1, 2, 3, 4, 5


class MATRIX_desc(C.Structure):
    """Byte-identical to the reference descriptor (include/descriptor.h:13-20)."""
    _fields_ = [("tile_size", C.c_int), ("matrix_nelements", C.c_long),
                ("tile_nelements", C.c_long), ("matrix_size", C.c_int),
                ("matrix", C.c_void_p)]


class Opts(C.Structure):
    _fields_ = [("device_matrix", C.c_int), ("precision", C.c_int), ("no_graph", C.c_int),
                ("no_lookahead", C.c_int), ("stream", C.c_void_p), ("refine", C.c_int),
                ("max_iter", C.c_int), ("tol", C.c_double), ("factor_tile", C.c_int), ("no_fallback", C.c_int),
                ("schedule", C.c_int), ("batch_grid", C.c_int), ("sweep_block", C.c_int),
                ("reserved", C.c_int * 3)]


class Report(C.Structure):
    _fields_ = [("factor_ms", C.c_double), ("solve_ms", C.c_double), ("h2d_ms", C.c_double),
                ("d2h_ms", C.c_double), ("gflop", C.c_double), ("tflops", C.c_double),
                ("backward_error", C.c_double), ("info", C.c_int), ("iterations", C.c_int),
                ("ntasks", C.c_long), ("nlaunches", C.c_long), ("gmres_iterations", C.c_int),
                ("precision_used", C.c_int), ("shifts", C.c_int), ("reserved", C.c_int * 5)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class PalialError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"palial error {code}: {msg}")
        self.code = code


_lib = None
_inited = False


def lib():
    """Load libpalial_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m palial_amd.build` first "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        dp = C.c_void_p
        L.palial_hip_init.argtypes = [C.c_int]
        L.palial_hip_set_option.argtypes = [C.c_char_p, C.c_int]
        L.palial_hip_strerror.restype = C.c_char_p
        L.palial_hip_strerror.argtypes = [C.c_int]
        L.palial_hip_last_error.restype = C.c_char_p
        L.palial_hip_generate.argtypes = [MATRIX_desc, C.c_int, C.c_uint64, C.c_double, C.POINTER(Opts)]
        L.palial_hip_cholesky.argtypes = [MATRIX_desc, C.POINTER(Opts), C.POINTER(Report)]
        L.palial_hip_lu.argtypes = [MATRIX_desc, dp, C.POINTER(Opts), C.POINTER(Report)]
        L.palial_hip_qr.argtypes = [MATRIX_desc, MATRIX_desc, C.POINTER(Opts), C.POINTER(Report)]
        L.palial_hip_dpotrf.argtypes = [C.c_int, dp, C.c_int, C.POINTER(Opts)]
        L.palial_hip_dgetrf.argtypes = [C.c_int, dp, C.c_int, C.POINTER(Opts)]
        L.palial_hip_dgeqrt.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp, dp, C.POINTER(Opts)]
        L.palial_hip_dormqr.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int,
                                        dp, C.c_int, C.POINTER(Opts)]
        L.palial_hip_dtsqrt.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, dp,
                                        C.POINTER(Opts)]
        L.palial_hip_dtsmqr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int,
                                        dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, C.POINTER(Opts)]
        L.palial_hip_dtsmqr_noalloc.argtypes = L.palial_hip_dtsmqr.argtypes
        L.palial_hip_qr_apply.argtypes = [C.c_int, C.c_int, MATRIX_desc, MATRIX_desc, dp, C.c_int, C.POINTER(Opts),
                                          C.POINTER(Report)]
        L.palial_hip_qr_solve.argtypes = [MATRIX_desc, MATRIX_desc, dp, C.c_int, C.c_int, C.POINTER(Opts),
                                          C.POINTER(Report)]
        L.palial_hip_transpose.argtypes = [MATRIX_desc, C.c_int, C.POINTER(Opts)]
        L.palial_hip_tile_task.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.POINTER(Opts)]
        L.palial_hip_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.palial_hip_free.argtypes = [dp]
        L.palial_hip_memcpy_h2d.argtypes = [dp, dp, C.c_size_t]
        L.palial_hip_memcpy_d2h.argtypes = [dp, dp, C.c_size_t]
        L.palial_hip_memcpy_d2d.argtypes = [dp, dp, C.c_size_t]
        L.palial_hip_copy.argtypes = [dp, dp, C.c_size_t]
        L.palial_hip_generate_kron.argtypes = [MATRIX_desc, dp, C.c_int, dp, C.c_int, C.POINTER(Opts)]
        L.palial_hip_solve.argtypes = [C.c_int, MATRIX_desc, dp, C.c_int, C.c_int, C.POINTER(Opts),
                                       C.POINTER(Report)]
        L.palial_hip_profile_stats.argtypes = [C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.palial_hip_update.argtypes = [C.c_int, MATRIX_desc, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.palial_hip_batched.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp, C.c_int, C.POINTER(Opts),
                                         C.POINTER(Report)]
        L.palial_hip_batched_generate.argtypes = [C.c_int, dp, dp, dp, dp, C.c_uint64, C.c_int, C.POINTER(Opts)]
        L.palial_hip_mlp_forward.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.POINTER(Opts)]
        L.palial_hip_mlp_train.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_int, C.c_float, dp, dp,
                                           C.POINTER(Opts)]
        L.palial_hip_trace.argtypes = [C.c_int]
        L.palial_hip_trace_records.restype = C.c_long
        L.palial_hip_trace_records.argtypes = [dp, C.c_long]
        L.palial_hip_task_name.restype = C.c_long
        L.palial_hip_task_name.argtypes = [C.c_int, C.c_int, C.c_long, C.c_char_p, C.c_int]
        L.palial_hip_profile_records.restype = C.c_long
        L.palial_hip_profile_records.argtypes = [dp, C.c_long]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        L = lib()
        raise PalialError(rc, f"{L.palial_hip_strerror(rc).decode()}: {L.palial_hip_last_error().decode()}")
    return rc


def init(device=0):
    global _inited
    if not _inited:
        _check(lib().palial_hip_init(device))
        _inited = True
        # release graphs, streams and device buffers while the HIP runtime is
        # still whole (before the shared libraries' own teardown at exit)
        atexit.register(finalize)


def finalize():
    global _inited
    if _inited:
        lib().palial_hip_finalize()
        _inited = False


def _opts(device=False, no_graph=False, no_lookahead=False, precision=PREC_FP64, stream=None,
          refine=0, max_iter=0, tol=0.0, factor_tile=0, no_fallback=False, schedule=0, batch_grid=0,
          sweep_block=0):
    o = Opts()
    o.schedule = schedule
    o.sweep_block = sweep_block
    o.batch_grid = batch_grid
    o.factor_tile = factor_tile
    o.no_fallback = int(bool(no_fallback))
    o.device_matrix = int(bool(device))
    o.no_graph = int(bool(no_graph))
    o.no_lookahead = int(bool(no_lookahead))
    o.precision = precision
    o.stream = stream
    o.refine = refine
    o.max_iter = max_iter
    o.tol = tol
    return o


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _desc_of(A, N=None, b=None):
    """Return (MATRIX_desc, on_device, keepalive)."""
    if isinstance(A, MATRIX_desc):
        return A, None, A
    if _is_torch(A):
        import torch
        if A.dtype != torch.float64 or not A.is_contiguous():
            raise TypeError("the matrix must be a contiguous float64 tensor in tile layout L")
        N = N or int(round(A.numel() ** 0.5))
        if N * N != A.numel() or not b:
            raise ValueError(f"a tensor of {A.numel()} elements is not N x N = {N} x {N}, or no tile size given")
        return MATRIX_desc(b, N * N, b * b, N, A.data_ptr()), A.is_cuda, A
    a = A
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("the matrix must be a C-contiguous float64 numpy array in tile layout L")
    N = N or int(round(a.size ** 0.5))
    if N * N != a.size or not b:
        raise ValueError(f"an array of {a.size} elements is not N x N = {N} x {N}, or no tile size given")
    return MATRIX_desc(b, N * N, b * b, N, a.ctypes.data), False, a


def _torch_stream(A):
    import torch
    return C.c_void_p(torch.cuda.current_stream(A.device).cuda_stream)


def generate(A, kind, seed=0, shift=0.0, N=None, b=None):
    """Fill A (layout L) with a seeded generator: GEN_SYM (S + shift I), GEN_DENSE
    (G + shift I), GEN_PALIAL (include/populate.h:35-66, seeded), GEN_IDENTITY."""
    init()
    d, dev, keep = _desc_of(A, N, b)
    o = _opts(device=dev)
    if dev:
        o.stream = _torch_stream(A)
    _check(lib().palial_hip_generate(d, kind, seed, float(shift), C.byref(o)))
    return A


def _run(fn, A, extra, N, b, **kw):
    init()
    d, dev, keep = _desc_of(A, N, b)
    o = _opts(device=dev if dev is not None else kw.pop("device", False), **kw)
    if dev:
        o.stream = _torch_stream(A)
    rep = Report()
    rc = _check(fn(d, *extra(dev), C.byref(o), C.byref(rep)))
    return rc, rep


def cholesky(A, b=None, N=None, **kw):
    """Upper tiled Cholesky in place.  Returns (info, report)."""
    return _run(lib().palial_hip_cholesky, A, lambda dev: (), N, b, **kw)


def lu(A, b=None, N=None, ipiv=None, **kw):
    """Tiled LU without pivoting in place.  Returns (info, report, ipiv)."""
    n = N or (A.matrix_size if isinstance(A, MATRIX_desc) else int(round(
        (A.numel() if _is_torch(A) else A.size) ** 0.5)))
    piv = np.zeros(n, dtype=np.int32) if ipiv is None else ipiv
    rc, rep = _run(lib().palial_hip_lu, A, lambda dev: (piv.ctypes.data,), N, b, **kw)
    return rc, rep, piv


def qr(A, S, b=None, N=None, **kw):
    """Flat-tree tiled QR in place; T factors go to S.  Returns (info, report)."""
    init()
    dA, dev, ka = _desc_of(A, N, b)
    dS, _, ks = _desc_of(S, N, b)
    o = _opts(device=dev, **kw)
    if dev:
        o.stream = _torch_stream(A)
    rep = Report()
    rc = _check(lib().palial_hip_qr(dA, dS, C.byref(o), C.byref(rep)))
    return rc, rep


def qr_apply(A, S, Cm, trans=TRANS, b=None, N=None, ncols=None, side=LEFT):
    """C <- Q^T C (trans=TRANS) or Q C (NOTRANS) with the Q of a finished qr(A, S).
    C: N x ncols in layout L (flat numpy array; ncols defaults to N), or a
    DeviceMatrix together with device A and S.  Returns the report."""
    init()
    if isinstance(A, DeviceMatrix):
        dA, dS, dev = A.desc, S.desc, True
        if isinstance(Cm, DeviceMatrix):
            cp, nc = Cm.ptr, Cm.N
        else:                       # a device pointer to N x ncols in layout L
            cp, nc = Cm, ncols
    else:
        dA, dev, _ = _desc_of(A, N, b)
        dS, _, _ = _desc_of(S, N, b)
        nc = ncols or (Cm.size // dA.matrix_size)
        cp = Cm.ctypes.data
    o = _opts(device=bool(dev))
    rep = Report()
    _check(lib().palial_hip_qr_apply(side, trans, dA, dS, cp, nc, C.byref(o), C.byref(rep)))
    return rep


def qr_solve(A, S, B, b=None, N=None):
    """X = R^-1 Q^T B from a finished qr(A, S) (host layout-L arrays); B is
    N x nrhs (nrhs <= 16).  Returns (X, report)."""
    init()
    dA, dev, _ = _desc_of(A, N, b)
    dS, _, _ = _desc_of(S, N, b)
    X = np.asfortranarray(np.array(B, dtype=np.float64, copy=True).reshape(dA.matrix_size, -1))
    o = _opts(device=bool(dev))
    rep = Report()
    _check(lib().palial_hip_qr_solve(dA, dS, X.ctypes.data, X.shape[1], X.shape[0], C.byref(o), C.byref(rep)))
    return X, rep


def transpose(A, average=False, b=None, N=None):
    """In-place A <- A^T, or A <- (A + A^T)/2 with average=True (layout L)."""
    init()
    if isinstance(A, DeviceMatrix):
        d, dev = A.desc, True
    else:
        d, dev, _ = _desc_of(A, N, b)
    _check(lib().palial_hip_transpose(d, int(bool(average)), C.byref(_opts(device=bool(dev)))))
    return A


# ------------------------------------------------------------------ tile kernels
def _f(a):
    a = np.asfortranarray(np.array(a, dtype=np.float64, copy=True))
    return a


def dpotrf(A):
    A = _f(A)
    init()
    info = _check(lib().palial_hip_dpotrf(A.shape[0], A.ctypes.data, A.shape[0], None))
    return A, info


def dgetrf(A):
    A = _f(A)
    init()
    _check(lib().palial_hip_dgetrf(A.shape[0], A.ctypes.data, A.shape[0], None))
    return A


def dgeqrt(A):
    A = _f(A)
    n = A.shape[0]
    T = np.zeros((n, n), order="F")
    tau = np.zeros(n)
    init()
    _check(lib().palial_hip_dgeqrt(n, n, A.ctypes.data, n, T.ctypes.data, n, tau.ctypes.data, None, None))
    return A, T, tau


def dormqr(V, T, Cm, side=LEFT, trans=TRANS):
    V, T, Cm = _f(V), _f(T), _f(Cm)
    n = V.shape[0]
    init()
    _check(lib().palial_hip_dormqr(side, trans, n, V.ctypes.data, n, T.ctypes.data, n, Cm.ctypes.data, n,
                                   None, n, None))
    return Cm


def dtsqrt(A1, A2):
    A1, A2 = _f(A1), _f(A2)
    n = A1.shape[0]
    T = np.zeros((n, n), order="F")
    tau = np.zeros(n)
    init()
    _check(lib().palial_hip_dtsqrt(n, A1.ctypes.data, n, A2.ctypes.data, n, T.ctypes.data, n,
                                   tau.ctypes.data, None, None))
    return A1, A2, T, tau


def dtsmqr(A1, A2, V, T, side=LEFT, trans=TRANS):
    A1, A2, V, T = _f(A1), _f(A2), _f(V), _f(T)
    n = A1.shape[0]
    init()
    _check(lib().palial_hip_dtsmqr(side, trans, n, n, n, n, A1.ctypes.data, n, A2.ctypes.data, n,
                                   V.ctypes.data, n, T.ctypes.data, n, None, n, None))
    return A1, A2


def tile_task(task, A, B, Cm):
    A, Cm = _f(A), _f(Cm)
    Bp = None
    if B is not None:
        B = _f(B)
        Bp = B.ctypes.data
    init()
    _check(lib().palial_hip_tile_task(task, A.shape[0], A.ctypes.data, Bp, Cm.ctypes.data, None))
    return Cm


# ------------------------------------------------------------------ device buffers
class DeviceMatrix:
    """An N x N layout-L matrix in device memory, owned by the library."""

    def __init__(self, N, b):
        init()
        self.N, self.b = N, b
        self.bytes = N * N * 8
        p = C.c_void_p()
        _check(lib().palial_hip_malloc(C.byref(p), self.bytes))
        self.ptr = p.value

    @property
    def desc(self):
        return MATRIX_desc(self.b, self.N * self.N, self.b * self.b, self.N, self.ptr)

    def generate(self, kind, seed=0, shift=0.0):
        o = _opts(device=True)
        _check(lib().palial_hip_generate(self.desc, kind, seed, float(shift), C.byref(o)))

    def generate_kron(self, A1, A2):
        A1 = np.asfortranarray(A1, dtype=np.float64)
        A2 = np.asfortranarray(A2, dtype=np.float64)
        o = _opts(device=True)
        _check(lib().palial_hip_generate_kron(self.desc, A1.ctypes.data, A1.shape[0], A2.ctypes.data, A2.shape[0],
                                              C.byref(o)))

    def copy_from(self, other):
        _check(lib().palial_hip_copy(self.ptr, other.ptr, self.bytes))

    def to_host(self):
        out = np.empty(self.N * self.N)
        _check(lib().palial_hip_memcpy_d2h(out.ctypes.data, self.ptr, self.bytes))
        return out

    def from_host(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().palial_hip_memcpy_h2d(self.ptr, a.ctypes.data, self.bytes))

    def free(self):
        if self.ptr:
            lib().palial_hip_free(self.ptr)
            self.ptr = None


def synchronize():
    _check(lib().palial_hip_synchronize())


def profile(enable):
    """0 off, 1 per-class totals, 2 also per-launch records (profile_records())."""
    init()
    _check(lib().palial_hip_profile(int(enable)))


KCLASS = {"gemm": 0, "diag": 1, "trsm": 2, "tsqrt": 3, "qrupd": 4, "pack": 5, "pkgemm": 6}


def profile_stats(cls):
    n = C.c_long(); ms = C.c_double(); fl = C.c_double(); by = C.c_double()
    _check(lib().palial_hip_profile_stats(KCLASS.get(cls, cls), C.byref(n), C.byref(ms), C.byref(fl),
                                          C.byref(by)))
    return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}


def update(algo, dm, k, precision=PREC_FP64, reps=1):
    """Apply the trailing update of step k to a DeviceMatrix reps times (test
    hook); returns the average time per rep in ms."""
    ms = C.c_double()
    _check(lib().palial_hip_update(0 if algo == "cholesky" else 1, dm.desc, k, precision, reps, C.byref(ms)))
    return ms.value


class Batch:
    """A batch of independent small matrices (orders n[i] <= 512; kind[i] 0 =
    Cholesky, 1 = LU without pivoting) stored back to back on the device,
    column-major, with one right-hand side each (palial_hip_batched)."""

    def __init__(self, n, kind, precision=PREC_FP64):
        init()
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.kind = np.ascontiguousarray(kind, dtype=np.int32)
        self.count = len(self.n)
        self.precision = precision
        self.dtype = np.float32 if precision == PREC_FP32 else np.float64
        sq = self.n.astype(np.int64) ** 2
        self.off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sq)[:-1]]), dtype=np.int64)
        self.boff = np.ascontiguousarray(np.concatenate([[0], np.cumsum(self.n.astype(np.int64))[:-1]]),
                                         dtype=np.int64)
        self.na, self.nb = int(sq.sum()), int(self.n.sum())
        es = np.dtype(self.dtype).itemsize
        pa, pb = C.c_void_p(), C.c_void_p()
        _check(lib().palial_hip_malloc(C.byref(pa), self.na * es))
        _check(lib().palial_hip_malloc(C.byref(pb), self.nb * es))
        self.A, self.B = pa.value, pb.value
        self.info = np.zeros(self.count, dtype=np.int32)

    def generate(self, seed=0):
        o = _opts(device=True)
        _check(lib().palial_hip_batched_generate(self.count, self.n.ctypes.data, self.off.ctypes.data,
                                                 self.kind.ctypes.data, self.A, seed, self.precision, C.byref(o)))

    def set(self, mats=None, rhs=None):
        """Upload matrices (list of n x n arrays) and / or right-hand sides (list of vectors)."""
        if mats is not None:
            flat = np.concatenate([np.asarray(m, dtype=self.dtype).ravel(order="F") for m in mats])
            _check(lib().palial_hip_memcpy_h2d(self.A, flat.ctypes.data, flat.nbytes))
        if rhs is not None:
            flat = np.concatenate([np.asarray(v, dtype=self.dtype).ravel() for v in rhs])
            _check(lib().palial_hip_memcpy_h2d(self.B, flat.ctypes.data, flat.nbytes))

    def get(self):
        """(list of n x n matrices, list of vectors) from the device."""
        a = np.empty(self.na, dtype=self.dtype)
        b = np.empty(self.nb, dtype=self.dtype)
        _check(lib().palial_hip_memcpy_d2h(a.ctypes.data, self.A, a.nbytes))
        _check(lib().palial_hip_memcpy_d2h(b.ctypes.data, self.B, b.nbytes))
        mats = [a[o:o + m * m].reshape(m, m, order="F") for o, m in zip(self.off, self.n)]
        vecs = [b[o:o + m] for o, m in zip(self.boff, self.n)]
        return mats, vecs

    def factor(self, solve=True, no_graph=False, grid=0):
        """Factor every matrix in place (and solve its right-hand side); returns the
        report.  One persistent launch of `grid` workgroups (0: what the CUs
        hold), replayed from a captured hipGraph unless no_graph."""
        o = _opts(device=True, no_graph=no_graph, batch_grid=grid)
        rep = Report()
        _check(lib().palial_hip_batched(self.count, self.n.ctypes.data, self.off.ctypes.data,
                                        self.kind.ctypes.data, self.A, self.B if solve else None,
                                        self.boff.ctypes.data if solve else None, self.info.ctypes.data,
                                        self.precision, C.byref(o), C.byref(rep)))
        return rep

    def free(self):
        for p in (self.A, self.B):
            if p:
                lib().palial_hip_free(p)
        self.A = self.B = None


def set_option(name, value):
    """Change a schedule setting that init reads from the environment ("pair":
    PALIAL_PAIR, "sweep_block": PALIAL_SWEEP_BLOCK); returns the previous value."""
    init()
    rc = lib().palial_hip_set_option(name.encode(), int(value))
    _check(rc)
    return rc


def trace_enable(on=True):
    """Per-task device timing of the following factorizations (palial_hip_trace)."""
    init()
    _check(lib().palial_hip_trace(int(bool(on))))


def trace_records():
    """The last traced factorization's tasks in creation order: (n, 6) array of
    (kind, k, m, n, device ms, start ms); see palial_amd.trace."""
    n = lib().palial_hip_trace_records(None, 0)
    if n < 0:
        raise PalialError(-3, "the task trace's timestamp buffer overflowed")
    out = np.zeros((n, 6))
    if n > 0:
        lib().palial_hip_trace_records(out.ctypes.data, n)
    return out


def task_names(algo, MT):
    """The library's reference task names in creation order (no GPU needed)."""
    code = {"cholesky": 0, "lu": 1, "qr": 2}[algo]
    n = lib().palial_hip_task_name(code, MT, 0, None, 0)
    buf = C.create_string_buffer(32)
    out = []
    for i in range(n):
        lib().palial_hip_task_name(code, MT, i, buf, 32)
        out.append(buf.value.decode())
    return out


def profile_records():
    """Per-launch records kept under profile(2): (n, 6) array of (class, b,
    precision, flops, bytes, ms)."""
    n = lib().palial_hip_profile_records(None, 0)
    out = np.zeros((max(n, 0), 6))
    if n > 0:
        lib().palial_hip_profile_records(out.ctypes.data, n)
    return out


def factor_device(algo, dm, S=None, **kw):
    """Factor a DeviceMatrix in place (device-resident, no host copies)."""
    o = _opts(device=True, **kw)
    rep = Report()
    if algo == "cholesky":
        rc = lib().palial_hip_cholesky(dm.desc, C.byref(o), C.byref(rep))
    elif algo == "lu":
        rc = lib().palial_hip_lu(dm.desc, None, C.byref(o), C.byref(rep))
    else:
        rc = lib().palial_hip_qr(dm.desc, S.desc, C.byref(o), C.byref(rep))
    _check(rc)
    return rc, rep


# ------------------------------------------------------------------ solve
def solve(A, B, algo="cholesky", precision=PREC_FP64, b=None, N=None, tol=0.0, max_iter=0, device=False,
          check=True, factor_tile=0, refine=0, no_fallback=False, sweep_block=0):
    """Solve A X = B with iterative refinement (no reference counterpart).

    A: layout-L matrix (numpy flat array, torch tensor or DeviceMatrix), left
    unmodified.  B: N x nrhs column-major right-hand sides (numpy, copied;
    or a device pointer with device=True).  precision: PREC_FP64, PREC_FP32,
    PREC_BF16, PREC_FP8 (E4M3 panels with a power-of-two scale per tile) or
    PREC_BF16X3 (split BF16); the last two need a factor tile that is a
    multiple of 128 and smaller than N.  sweep_block: tiles per step of the
    correction sweeps (0 automatic, 1 one tile).  Returns (X, report, rc)."""
    init()
    if isinstance(A, DeviceMatrix):
        d, dev = A.desc, True
    else:
        d, dev, keep = _desc_of(A, N, b)
    o = _opts(device=bool(dev), precision=precision, tol=tol, max_iter=max_iter, factor_tile=factor_tile,
              refine=refine, no_fallback=no_fallback, sweep_block=sweep_block)
    rep = Report()
    code = {"cholesky": 0, "lu": 1}[algo]
    if dev and not isinstance(B, np.ndarray):
        ptr, nrhs, ldb = B
        rc = lib().palial_hip_solve(code, d, ptr, nrhs, ldb, C.byref(o), C.byref(rep))
        X = None
    else:
        X = np.asfortranarray(np.array(B, dtype=np.float64, copy=True).reshape(d.matrix_size, -1))
        if dev:
            raise ValueError("host right-hand sides need a host matrix")
        rc = lib().palial_hip_solve(code, d, X.ctypes.data, X.shape[1], X.shape[0], C.byref(o), C.byref(rep))
    if check and rc < 0:
        _check(rc)
    return X, rep, rc


class CondGenerator:
    """Device matrices A = Q diag(s) Q^T with Q the orthogonal factor of the
    library's own tiled Householder QR of a seeded N x N standard normal
    matrix (GEN_GAUSS) and s_i = kappa^(-i/(N-1)) (GEN_GEOM_DIAG), so that the
    2-norm condition number is kappa.  The QR is computed once per generator
    and reused for every kappa:
        A <- diag(s);  A <- Q A;  A <- A^T;  A <- Q A;  A <- (A + A^T)/2."""

    def __init__(self, N, b, seed=0):
        self.N, self.b = N, b
        self.G = DeviceMatrix(N, b)
        self.S = DeviceMatrix(N, b)
        self.G.generate(GEN_GAUSS, seed=seed)
        factor_device("qr", self.G, self.S)

    def fill(self, A, kappa):
        """Write Q diag(s) Q^T for this kappa into DeviceMatrix A (same N, b)."""
        A.generate(GEN_GEOM_DIAG, shift=float(kappa))
        qr_apply(self.G, self.S, A, trans=NOTRANS)
        transpose(A)
        qr_apply(self.G, self.S, A, trans=NOTRANS)
        transpose(A, average=True)
        return A

    def free(self):
        self.G.free()
        self.S.free()


def kron_spd_factors(N, kappa, seed=0, n1=None):
    """Factors of A = Q diag(s) Q^T with Q = Q1 (x) Q2 (Householder QR of seeded
    Gaussians) and s geometric from 1 to 1/kappa: returns (A1, A2) with
    A = A1 (x) A2.  s_k = r^k, k = k1*n2 + k2, splits as r^(n2 k1) * r^k2."""
    if n1 is None:
        n1 = 1 << ((N.bit_length() - 1) // 2)
        while N % n1:
            n1 //= 2
    n2 = N // n1
    rng = np.random.default_rng(seed)
    Q1, _ = np.linalg.qr(rng.standard_normal((n1, n1)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n2, n2)))
    r = kappa ** (-1.0 / (N - 1)) if N > 1 else 1.0
    d1 = r ** (n2 * np.arange(n1))
    d2 = r ** np.arange(n2)
    A1 = (Q1 * d1) @ Q1.T
    A2 = (Q2 * d2) @ Q2.T
    return (A1 + A1.T) / 2, (A2 + A2.T) / 2
