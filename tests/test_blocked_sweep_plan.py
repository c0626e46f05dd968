"""CPU tests of the blocked correction sweeps (no reference counterpart: the
reference only factors).  The library's own host views of what a solve
launches -- the sweep launch order, each launch's tile map, the gather map
and the GEMM launches that build the q x q block inverses, and the GEMM task
decode -- are executed here in numpy on a small factored matrix whose
unread tiles hold NaN.  The solution must equal numpy's dense solve, and the
block inverses the exact inverses of the diagonal blocks, so a wrong tile,
transpose, level order or operand shows up as a NaN or a wrong digit."""
import ctypes as C

import numpy as np
import pytest

GV_DIAG = 0
GM_BI_W, GM_BI_X = 21, 22


def lib():
    import palial_amd
    L = palial_amd.lib()
    L.palial_hip_sweep_plan.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int]
    L.palial_hip_gemv_tiles.argtypes = [C.c_int] * 7 + [C.POINTER(C.c_long), C.c_long]
    L.palial_hip_gather_map.argtypes = [C.c_int] * 5 + [C.POINTER(C.c_long), C.c_long]
    L.palial_hip_blkinv_plan.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int]
    L.palial_hip_blkinv_tiles.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_long), C.c_long]
    return L


def sweep_plan(algo, MT, q):
    buf = (C.c_int * (7 * 4 * MT))()
    n = lib().palial_hip_sweep_plan(algo, MT, q, buf, 4 * MT)
    assert n > 0
    return np.array(buf[:7 * n]).reshape(n, 7)


def gemv_tiles(mode, trans, MT, b, q, k, ntasks):
    buf = (C.c_long * max(1, ntasks * (3 + q)))()
    n = lib().palial_hip_gemv_tiles(mode, trans, MT, b, q, k, ntasks, buf, len(buf))
    assert n == ntasks
    return np.array(buf[:ntasks * (3 + q)]).reshape(ntasks, 3 + q)


def gather_map(MT, b, q, upper, dtrans):
    buf = (C.c_long * (4 * MT * q))()
    n = lib().palial_hip_gather_map(MT, b, q, upper, dtrans, buf, len(buf))
    assert n == MT * q
    return np.array(buf[:4 * n]).reshape(n, 4)


def blkinv_plan(q):
    buf = (C.c_double * (5 * q * q))()
    n = lib().palial_hip_blkinv_plan(q, buf, q * q)
    assert n >= 0
    return np.array(buf[:5 * n]).reshape(n, 5)


def blkinv_tiles(mode, MT, b, q, bd, bt):
    ntasks = MT // q * (q - bd)
    buf = (C.c_long * (8 * ntasks))()
    n = lib().palial_hip_blkinv_tiles(mode, MT, b, q, bd, bt, buf, len(buf))
    assert n == ntasks
    return np.array(buf[:8 * n]).reshape(n, 8)


def tile_off(b, MT, m, n):
    return b * b * (m + MT * n)


def get(arr, off, b):
    return arr[off:off + b * b].reshape(b, b, order="F")


def put(arr, off, b, T):
    arr[off:off + b * b] = T.ravel(order="F")


def factor(algo, MT, b, seed):
    """Dense A, its factor in layout L with every tile the sweeps must not
    read set to NaN, and the tile inverses the solve computes (Cholesky:
    U_kk^-T; LU: U_kk^-1 and L_kk^-1)."""
    N = MT * b
    r = np.random.default_rng(seed)
    G = r.standard_normal((N, N))
    if algo == 0:
        A = G @ G.T / N + np.eye(N)
        U = np.linalg.cholesky(A).T
        Lm = None
    else:
        A = G / np.sqrt(N) + 4 * np.eye(N)
        Lm, U = np.eye(N), A.copy()
        for j in range(N):
            Lm[j + 1:, j] = U[j + 1:, j] / U[j, j]
            U[j + 1:, :] -= np.outer(Lm[j + 1:, j], U[j, :])
        U = np.triu(U)
    F = np.full(N * N, np.nan)
    for m in range(MT):
        for n in range(MT):
            if m < n:
                put(F, tile_off(b, MT, m, n), b, U[m * b:(m + 1) * b, n * b:(n + 1) * b])
            elif m > n and algo == 1:
                put(F, tile_off(b, MT, m, n), b, Lm[m * b:(m + 1) * b, n * b:(n + 1) * b])
    sl = [slice(k * b, (k + 1) * b) for k in range(MT)]
    if algo == 0:
        DU = np.concatenate([np.linalg.inv(U[s, s]).T.ravel(order="F") for s in sl])
        DL = None
    else:
        DU = np.concatenate([np.linalg.inv(U[s, s]).ravel(order="F") for s in sl])
        DL = np.concatenate([np.linalg.inv(Lm[s, s]).ravel(order="F") for s in sl])
    return A, F, U, Lm, DU, DL


def block_inverse(F, MT, b, q, upper, dtrans, D):
    """Executes the gather map and the GEMM launches of the block inverses."""
    X = np.full(q * MT * b * b, np.nan)
    W = [X, np.full(q * MT * b * b, np.nan), np.full(max(1, (q - 1) * (MT // q)) * b * b, np.nan)]
    for kind, tr, src, dst in gather_map(MT, b, q, upper, dtrans):
        if kind == 0:
            put(X, dst, b, np.zeros((b, b)))
        else:
            T = get(D if kind == 1 else F, src, b)
            put(X if kind == 1 else W[1], dst, b, T.T if tr else T)
    for mode, bd, bt, alpha, beta in blkinv_plan(q):
        for ba, oa, bb_, ob, bc, oc, bdd, od in blkinv_tiles(int(mode), MT, b, q, int(bd), int(bt)):
            acc = alpha * (get(W[ba], oa, b) @ get(W[bb_], ob, b))
            if bdd >= 0:
                acc = acc + beta * get(W[bdd], od, b)
            put(W[bc], oc, b, acc)
    return X


def run_sweeps(algo, MT, b, q, F, invs, rhs):
    N = MT * b
    vec = [rhs.copy(), np.full(N, np.nan), np.full(N, np.nan)]
    for mode, k, trans, ntasks, inv, vin, vout in sweep_plan(algo, MT, q):
        for row in gemv_tiles(mode, trans, MT, b, q, k, ntasks):
            o, j0, j1 = row[:3]
            acc = np.zeros(b)
            for j in range(j0, j1):
                T = get(invs[inv] if mode == GV_DIAG else F, row[3 + j], b)
                seg = vec[vin][(k * q + j) * b:(k * q + j + 1) * b]
                acc += (T.T if trans else T) @ seg
            if mode == GV_DIAG:
                vec[vout][o * b:(o + 1) * b] = acc
            else:
                vec[vout][o * b:(o + 1) * b] -= acc
    return vec[2]


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("MT,q", [(4, 1), (4, 2), (6, 1), (6, 2), (6, 3), (8, 2), (8, 4)])
def test_blocked_sweeps_solve(algo, MT, q):
    b = 8
    A, F, U, Lm, DU, DL = factor(algo, MT, b, seed=100 * algo + 10 * MT + q)
    N = MT * b
    if q == 1:
        invs = [DU, DL]
    else:
        XU = block_inverse(F, MT, b, q, 1, algo, DU)
        XL = block_inverse(F, MT, b, q, 0, 0, DL) if algo == 1 else None
        for K in range(MT // q):
            s = slice(K * q * b, (K + 1) * q * b)
            blk = np.block([[get(XU, ((K * q + i) * q + j) * b * b, b) for j in range(q)] for i in range(q)])
            assert np.allclose(blk, np.linalg.inv(U[s, s].T), rtol=0, atol=1e-12), K
            if algo == 1:
                blk = np.block([[get(XL, ((K * q + i) * q + j) * b * b, b) for j in range(q)] for i in range(q)])
                assert np.allclose(blk, np.linalg.inv(Lm[s, s]), rtol=0, atol=1e-12), K
        invs = [XU, XL]
    rhs = np.random.default_rng(7).standard_normal(N)
    x = run_sweeps(algo, MT, b, q, F, invs, rhs)
    ref = np.linalg.solve(A, rhs)
    assert np.all(np.isfinite(x))
    assert np.abs(x - ref).max() <= 1e-10 * np.abs(ref).max()


def test_plan_shapes():
    """q tiles per step: 2 MT/q diagonal launches and 2 (MT/q - 1) updates;
    the block inverses take sum_{d=1}^{q-1} (d + 1) launches."""
    for algo in (0, 1):
        for MT, q in ((8, 1), (8, 2), (8, 4), (6, 3)):
            p = sweep_plan(algo, MT, q)
            assert len(p) == 2 * (MT // q) + 2 * (MT // q - 1)
            assert (p[:, 0] == GV_DIAG).sum() == 2 * (MT // q)
            assert p[p[:, 0] == GV_DIAG][:, 3].tolist() == [q] * (2 * (MT // q))
    for q in (1, 2, 3, 4, 8):
        assert len(blkinv_plan(q)) == sum(d + 1 for d in range(1, q))
    L = lib()
    buf = (C.c_int * 64)()
    assert L.palial_hip_sweep_plan(0, 6, 4, buf, 8) < 0        # q must divide MT
    assert L.palial_hip_sweep_plan(0, 8, 2, buf, 3) < 0        # cap too small


@pytest.mark.parametrize("algo", [0, 1])
def test_blocked_sweeps_fp32_accuracy(algo):
    """The mixed tiers sweep FP32 copies of the factor and of the inverses
    (built in FP64, then rounded), with FP64 sums.  Executing the library's
    plans that way on an ill-conditioned factor (Cholesky of an SPD matrix with
    kappa 1e8; LU with a non-normal unit L of condition ~1e6 and an ill-scaled
    U), blocks of q tiles are as accurate as one tile per step: the rounding
    of a block inverse costs no more than the roundings of the tile steps it
    replaces."""
    MT, b = 16, 16
    N = MT * b
    r = np.random.default_rng(31)
    if algo == 0:
        Q, _ = np.linalg.qr(r.standard_normal((N, N)))
        A = (Q * np.geomspace(1.0, 1e-8, N)) @ Q.T
        A = (A + A.T) / 2
        U = np.linalg.cholesky(A).T
        Lm = None
    else:
        Lm = np.eye(N) + 0.1 * np.tril(r.standard_normal((N, N)), -1)
        U = np.diag(np.geomspace(1.0, 1e-6, N)[r.permutation(N)]) @ (np.eye(N) + np.triu(r.standard_normal((N, N)) * 0.05, 1))
        A = Lm @ U
    F = np.full(N * N, np.nan)
    for m in range(MT):
        for n in range(MT):
            if m < n:
                put(F, tile_off(b, MT, m, n), b, U[m * b:(m + 1) * b, n * b:(n + 1) * b])
            elif m > n and algo == 1:
                put(F, tile_off(b, MT, m, n), b, Lm[m * b:(m + 1) * b, n * b:(n + 1) * b])
    sl = [slice(k * b, (k + 1) * b) for k in range(MT)]
    if algo == 0:
        DU = np.concatenate([np.linalg.inv(U[s, s]).T.ravel(order="F") for s in sl])
        DL = None
    else:
        DU = np.concatenate([np.linalg.inv(U[s, s]).ravel(order="F") for s in sl])
        DL = np.concatenate([np.linalg.inv(Lm[s, s]).ravel(order="F") for s in sl])
    f32 = lambda v: None if v is None else v.astype(np.float32).astype(np.float64)
    rhs = r.standard_normal(N)
    exact = np.linalg.solve(A, rhs)
    errs = {}
    for q in (1, 2, 4):
        if q == 1:
            invs = [f32(DU), f32(DL)]
        else:
            XU = block_inverse(F, MT, b, q, 1, algo, DU)
            XL = block_inverse(F, MT, b, q, 0, 0, DL) if algo == 1 else None
            invs = [f32(XU), f32(XL)]
        x = run_sweeps(algo, MT, b, q, f32(F), invs, rhs)
        errs[q] = np.linalg.norm(x - exact) / np.linalg.norm(exact)
    assert errs[1] < 1e-3, errs                  # the FP32 sweep itself is meaningful here
    assert errs[2] <= 1.5 * errs[1] and errs[4] <= 1.5 * errs[1], errs


def test_random_blocked_sweeps():
    """Random tile counts, tile sizes and step widths (every q that divides
    the tile count twice over): the library's plans still solve exactly."""
    from hypothesis import assume, given, settings, strategies as st

    @settings(max_examples=30, deadline=None, derandomize=True)
    @given(algo=st.sampled_from([0, 1]), MT=st.integers(2, 12), q=st.integers(1, 6), b=st.sampled_from([4, 8]))
    def check(algo, MT, q, b):
        assume(MT % q == 0 and MT >= 2 * q)
        A, F, U, Lm, DU, DL = factor(algo, MT, b, seed=MT * 31 + q * 7 + b + algo)
        if q == 1:
            invs = [DU, DL]
        else:
            invs = [block_inverse(F, MT, b, q, 1, algo, DU),
                    block_inverse(F, MT, b, q, 0, 0, DL) if algo == 1 else None]
        rhs = np.random.default_rng(MT + q).standard_normal(MT * b)
        x = run_sweeps(algo, MT, b, q, F, invs, rhs)
        ref = np.linalg.solve(A, rhs)
        assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()

    check()
