"""Python view of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and bench.py's CPU-baseline leg import this
package.  It wraps ``oracle/liboracle.so`` (built by ``oracle/Makefile``) with
ctypes and numpy; arrays in layout L are flat float64 arrays of N*N elements
whose tile (m, n) starts at ``b*b*(m + (N//b)*n)`` (include/tile_address.h:15 of
the reference) and is column-major inside with ld = b.
"""
import ctypes as C
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

LEFT, RIGHT = 141, 142
NOTRANS, TRANS = 111, 112


class Desc(C.Structure):
    """Same layout as the reference MATRIX_desc (include/descriptor.h:13-20)."""
    _fields_ = [("tile_size", C.c_int), ("matrix_nelements", C.c_long),
                ("tile_nelements", C.c_long), ("matrix_size", C.c_int),
                ("matrix", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.orc_uniform.restype = C.c_double
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        for f in ("orc_gen_sym_shift", "orc_gen_dense_shift"):
            getattr(L, f).argtypes = [Desc, C.c_uint64, C.c_double]
        L.orc_gen_palial.argtypes = [Desc, C.c_uint64]
        L.orc_gen_identity.argtypes = [Desc]
        L.orc_gen_gauss.argtypes = [Desc, C.c_uint64]
        L.orc_gen_geom_diag.argtypes = [Desc, C.c_double]
        L.orc_gauss.restype = C.c_double
        L.orc_gauss.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_gen_special4x4.argtypes = [Desc]
        L.orc_use_openblas.argtypes = [C.c_char_p]
        L.orc_wtime.restype = C.c_double
        L.orc_tile_offset.restype = C.c_long
        L.orc_tile_offset.argtypes = [C.c_int] * 4
        L.orc_elem_offset.restype = C.c_long
        L.orc_elem_offset.argtypes = [C.c_int] * 4
        L.orc_cholesky.argtypes = [Desc]
        L.orc_lu.argtypes = [Desc]
        L.orc_qr.argtypes = [Desc, Desc]
        L.orc_cholesky_omp.argtypes = [Desc, C.c_int]
        L.orc_lu_omp.argtypes = [Desc, C.c_int]
        L.orc_qr_omp.argtypes = [Desc, Desc, C.c_int]
        L.orc_task_names.restype = C.c_long
        L.orc_task_names.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_size_t]
        L.orc_chol_solve.argtypes = [Desc, dp, C.c_int, C.c_int]
        L.orc_lu_solve.argtypes = [Desc, dp, C.c_int, C.c_int]
        L.orc_dpotrf_tile.argtypes = [C.c_int, dp, C.c_int]
        L.orc_dgetrf.argtypes = [C.c_int, dp, C.c_int]
        L.orc_dgeqrt.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp, dp]
        L.orc_dormqr.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp,
                                 C.c_int, dp, C.c_int]
        L.orc_dtsqrt.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, dp]
        L.orc_dtsmqr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp,
                                 C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int, dp, C.c_int]
        for f in ("orc_chol_trsm", "orc_chol_syrk", "orc_lu_trsm_r", "orc_lu_trsm_l"):
            getattr(L, f).argtypes = [C.c_int, dp, dp]
        for f in ("orc_chol_gemm", "orc_lu_gemm"):
            getattr(L, f).argtypes = [C.c_int, dp, dp, dp]
        _LIB = L
    return _LIB


def ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def desc(a, N, b):
    assert a.size == N * N
    return Desc(b, N * N, b * b, N, a.ctypes.data)


# ---------------------------------------------------------------- backends
def openblas_path():
    cands = glob.glob("/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so")
    try:
        import scipy
        cands += glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs",
                                        "libscipy_openblas*.so"))
    except ImportError:
        pass
    return cands[0] if cands else None


def use_openblas():
    p = openblas_path()
    return bool(p) and bool(lib().orc_use_openblas(p.encode()))


def use_naive():
    lib().orc_use_naive()


def openblas_threads(n):
    lib().orc_openblas_threads(n)


# ---------------------------------------------------------------- layout
def to_dense(t, N, b):
    """layout L (flat) -> dense N x N (row index i, column index j)."""
    MT = N // b
    return np.ascontiguousarray(t.reshape(MT, MT, b, b).transpose(1, 3, 0, 2).reshape(N, N))


def from_dense(g, N, b):
    MT = N // b
    return np.ascontiguousarray(g.reshape(MT, b, MT, b).transpose(2, 0, 3, 1).reshape(N * N))


# ---------------------------------------------------------------- generators
def gen(kind, N, b, seed=0, shift=None):
    a = np.empty(N * N, dtype=np.float64)
    d = desc(a, N, b)
    L = lib()
    if kind == "sym":
        L.orc_gen_sym_shift(d, seed, float(N if shift is None else shift))
    elif kind == "dense":
        L.orc_gen_dense_shift(d, seed, float(0.0 if shift is None else shift))
    elif kind == "palial":
        L.orc_gen_palial(d, seed)
    elif kind == "identity":
        L.orc_gen_identity(d)
    elif kind == "gauss":
        L.orc_gen_gauss(d, seed)
    elif kind == "geomdiag":
        L.orc_gen_geom_diag(d, float(shift))
    elif kind == "special4x4":
        assert N == 4 and b == 2
        L.orc_gen_special4x4(d)
    else:
        raise ValueError(kind)
    return a


def uniform(seed, i, j):
    return lib().orc_uniform(seed, i, j)


# ---------------------------------------------------------------- algorithms
def cholesky(a, N, b, omp_threads=0):
    a = np.array(a, dtype=np.float64, copy=True)
    if omp_threads:
        lib().orc_cholesky_omp(desc(a, N, b), omp_threads)
        return a, 0
    info = lib().orc_cholesky(desc(a, N, b))
    return a, info


def lu(a, N, b, omp_threads=0):
    a = np.array(a, dtype=np.float64, copy=True)
    if omp_threads:
        lib().orc_lu_omp(desc(a, N, b), omp_threads)
    else:
        lib().orc_lu(desc(a, N, b))
    return a


def qr(a, N, b, omp_threads=0):
    a = np.array(a, dtype=np.float64, copy=True)
    s = np.zeros(N * N, dtype=np.float64)
    if omp_threads:
        lib().orc_qr_omp(desc(a, N, b), desc(s, N, b), omp_threads)
    else:
        lib().orc_qr(desc(a, N, b), desc(s, N, b))
    return a, s


def chol_solve(u, N, b, rhs):
    x = np.asfortranarray(np.array(rhs, dtype=np.float64, copy=True).reshape(N, -1))
    lib().orc_chol_solve(desc(u, N, b), x.ctypes.data_as(C.POINTER(C.c_double)), x.shape[1], N)
    return x


def lu_solve(f, N, b, rhs):
    x = np.asfortranarray(np.array(rhs, dtype=np.float64, copy=True).reshape(N, -1))
    lib().orc_lu_solve(desc(f, N, b), x.ctypes.data_as(C.POINTER(C.c_double)), x.shape[1], N)
    return x


def task_names(algo, MT):
    code = {"cholesky": 0, "lu": 1, "qr": 2}[algo]
    n = lib().orc_task_names(code, MT, None, 0)
    buf = C.create_string_buffer(40 * n + 64)
    lib().orc_task_names(code, MT, buf, len(buf))
    return buf.value.decode().splitlines()


# ---------------------------------------------------------------- tile kernels
def tile_potrf(A):
    A = np.asfortranarray(A.copy())
    info = lib().orc_dpotrf_tile(A.shape[0], A.ctypes.data_as(C.POINTER(C.c_double)), A.shape[0])
    return A, info


def _fp(A):
    return A.ctypes.data_as(C.POINTER(C.c_double))


def tile_getrf(A):
    A = np.asfortranarray(A.copy())
    lib().orc_dgetrf(A.shape[0], _fp(A), A.shape[0])
    return A


def tile_geqrt(A):
    n = A.shape[0]
    A = np.asfortranarray(A.copy())
    T = np.zeros((n, n), order="F")
    tau = np.zeros(n)
    work = np.zeros(n * n)
    lib().orc_dgeqrt(n, n, _fp(A), n, _fp(T), n, _fp(tau), _fp(work))
    return A, T, tau


def tile_ormqr(V, T, Cm):
    n = V.shape[0]
    V = np.asfortranarray(V); T = np.asfortranarray(T)
    Cm = np.asfortranarray(Cm.copy())
    work = np.zeros(n * n)
    r = lib().orc_dormqr(LEFT, TRANS, n, _fp(V), n, _fp(T), n, _fp(Cm), n, _fp(work), n)
    assert r == 0
    return Cm


def tile_tsqrt(A1, A2):
    n = A1.shape[0]
    A1 = np.asfortranarray(A1.copy()); A2 = np.asfortranarray(A2.copy())
    T = np.zeros((n, n), order="F")
    tau = np.zeros(n)
    work = np.zeros(n * n)
    lib().orc_dtsqrt(n, _fp(A1), n, _fp(A2), n, _fp(T), n, _fp(tau), _fp(work))
    return A1, A2, T, tau


def tile_tsmqr(A1, A2, V, T):
    n = A1.shape[0]
    A1 = np.asfortranarray(A1.copy()); A2 = np.asfortranarray(A2.copy())
    V = np.asfortranarray(V); T = np.asfortranarray(T)
    work = np.zeros(n * n)
    r = lib().orc_dtsmqr(LEFT, TRANS, n, n, n, n, _fp(A1), n, _fp(A2), n, _fp(V), n, _fp(T), n,
                         _fp(work), n)
    assert r == 0
    return A1, A2


def _tile_task(fname, *tiles):
    out = np.asfortranarray(tiles[-1].copy())
    ins = [np.asfortranarray(t) for t in tiles[:-1]]
    getattr(lib(), fname)(out.shape[0], *[_fp(t) for t in ins], _fp(out))
    return out


def chol_trsm(Akk, Akm):
    return _tile_task("orc_chol_trsm", Akk, Akm)


def chol_syrk(Akm, Amm):
    return _tile_task("orc_chol_syrk", Akm, Amm)


def chol_gemm(Akn, Akm, Anm):
    return _tile_task("orc_chol_gemm", Akn, Akm, Anm)


def lu_trsm_r(Akk, Amk):
    return _tile_task("orc_lu_trsm_r", Akk, Amk)


def lu_trsm_l(Akk, Akn):
    return _tile_task("orc_lu_trsm_l", Akk, Akn)


def lu_gemm(Amk, Akn, Amn):
    return _tile_task("orc_lu_gemm", Amk, Akn, Amn)


# ---------------------------------------------------------------- residuals
EPS = np.finfo(np.float64).eps


def resid_chol(A, Ufac):
    """||A - U^T U||_F / (n eps ||A||_F) on dense inputs (U = upper triangle of Ufac)."""
    U = np.triu(Ufac)
    n = A.shape[0]
    return np.linalg.norm(A - U.T @ U) / (n * EPS * np.linalg.norm(A))


def resid_lu(A, F):
    n = A.shape[0]
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    return np.linalg.norm(A - L @ U) / (n * EPS * np.linalg.norm(A))


def gauss(seed, i, j):
    return lib().orc_gauss(seed, i, j)


def qr_blocks(F, S, N, b):
    """The block reflectors of a flat-tree tiled QR (src/qr.h:18-190) in the
    order the factorization applied them: (rows, Y, T) with the transformation
    C[rows] <- (I - Y T Y^T)^T C[rows]; Y = unit lower V of tile (k,k) for the
    GEQRT block of step k, Y = [I; A(m,k)] for the TS block (k, m)."""
    MT = N // b

    def tile(M, m, n):
        o = b * b * (m + MT * n)
        return M[o:o + b * b].reshape(b, b).T

    ops = []
    for k in range(MT):
        ops.append(([k], np.tril(tile(F, k, k), -1) + np.eye(b), np.triu(tile(S, k, k))))
        for m in range(k + 1, MT):
            ops.append(([k, m], np.vstack([np.eye(b), tile(F, m, k)]), np.triu(tile(S, m, k))))
    return ops


def apply_q(F, S, N, b, Cd, trans):
    """Q^T C (trans) or Q C of a dense N x n matrix Cd, Q = H_1 H_2 ... with
    H = I - Y T Y^T the blocks of qr_blocks."""
    X = np.array(Cd, dtype=np.float64, copy=True)
    ops = qr_blocks(F, S, N, b)
    for rows, Y, T in (ops if trans else reversed(ops)):
        idx = np.concatenate([np.arange(r * b, (r + 1) * b) for r in rows])
        Z = X[idx, :]
        X[idx, :] = Z - Y @ ((T.T if trans else T) @ (Y.T @ Z))
    return X


def form_q(F, S, N, b):
    return apply_q(F, S, N, b, np.eye(N), trans=False)


def resid_qr(A, F, S, N, b):
    """Scaled residuals (||A - QR||_F / (n eps ||A||_F), ||I - Q^T Q||_F / (n eps))
    of a flat-tree tiled QR (src/qr.h:18-190) from its output in layout L:
    F holds R (upper part), the GEQRT reflectors V (unit lower, tile (k,k))
    and the TS reflectors (tile (m,k), m > k); S the T factors.  The
    factorization applied C <- (I - Y T Y^T)^T C per step (dlarfb 'L','T';
    src/srcqr/dtsmqr.h), so Q = H_1 H_2 ... with H = I - Y T Y^T, formed here
    by applying the H in reverse order to the identity."""
    Q = form_q(F, S, N, b)
    R = np.triu(to_dense(F, N, b))
    r1 = np.linalg.norm(A - Q @ R) / (N * EPS * np.linalg.norm(A))
    r2 = np.linalg.norm(np.eye(N) - Q.T @ Q) / (N * EPS)
    return r1, r2
