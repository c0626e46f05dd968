"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerance (north star): factors within 1e-10 relative (max-abs difference over
max-abs reference value) of the oracle's sequential program order on the same
seeded inputs; scaled residuals below 10; generators and pivots bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(x, ref):
    x = np.asarray(x); ref = np.asarray(ref)
    den = np.abs(ref).max()
    return np.abs(x - ref).max() / (den if den > 0 else 1.0)


def rand_tile(n, seed, diag=0.0):
    r = np.random.default_rng(seed)
    a = r.uniform(-0.5, 0.5, (n, n))
    return np.asfortranarray(a + diag * np.eye(n))


def spd_tile(n, seed):
    a = rand_tile(n, seed)
    return np.asfortranarray((a + a.T) / 2 + n * np.eye(n))


TILE_SIZES = [4, 32, 64, 96, 128, 256, 512]


# ------------------------------------------------------------------ generators
@pytest.mark.parametrize("kind,N,b", [("sym", 256, 64), ("dense", 512, 128), ("palial", 96, 32),
                                      ("identity", 64, 16), ("sym", 2048, 256)])
def test_generator_bit_exact(hip, orc, kind, N, b):
    code = {"sym": hip.GEN_SYM, "dense": hip.GEN_DENSE, "palial": hip.GEN_PALIAL, "identity": hip.GEN_IDENTITY}[kind]
    shift = float(N) if kind in ("sym", "dense") else 0.0
    a = np.zeros(N * N)
    hip.generate(a, code, seed=7, shift=shift, N=N, b=b)
    ref = orc.gen(kind, N, b, seed=7, shift=shift) if kind in ("sym", "dense") else orc.gen(kind, N, b, seed=7)
    assert np.array_equal(a.view(np.uint64), ref.view(np.uint64))


# ------------------------------------------------------------------ BLAS tile tasks
@pytest.mark.parametrize("n", TILE_SIZES)
def test_tile_tasks(hip, orc, n):
    U, _ = orc.tile_potrf(spd_tile(n, 1))
    U = np.triu(U) + np.tril(rand_tile(n, 9), -1)   # strict lower holds junk, never read
    L = orc.tile_getrf(rand_tile(n, 2, diag=n))
    A, B, Cm = rand_tile(n, 3), rand_tile(n, 4), rand_tile(n, 5)
    cases = [
        (hip.TASK_CHOL_TRSM, U, None, Cm, orc.chol_trsm(U, Cm)),
        (hip.TASK_CHOL_SYRK, A, None, Cm, orc.chol_syrk(A, Cm)),
        (hip.TASK_CHOL_GEMM, A, B, Cm, orc.chol_gemm(A, B, Cm)),
        (hip.TASK_LU_TRSM_R, L, None, Cm, orc.lu_trsm_r(L, Cm)),
        (hip.TASK_LU_TRSM_L, L, None, Cm, orc.lu_trsm_l(L, Cm)),
        (hip.TASK_LU_GEMM, A, B, Cm, orc.lu_gemm(A, B, Cm)),
    ]
    for task, a, b, c, ref in cases:
        out = hip.tile_task(task, a, b, c)
        assert rel(out, ref) < TOL, (task, rel(out, ref))
        if task == hip.TASK_CHOL_SYRK:   # strictly lower part untouched
            assert np.array_equal(np.tril(out, -1), np.tril(c, -1))


# ------------------------------------------------------------------ panel kernels
@pytest.mark.parametrize("n", TILE_SIZES)
def test_dpotrf_tile(hip, orc, n):
    a = spd_tile(n, 11)
    junk = np.tril(rand_tile(n, 12), -1)
    a = np.triu(a) + junk   # dpotrf('U') reads and writes the upper triangle only
    out, info = hip.dpotrf(a)
    ref, rinfo = orc.tile_potrf(a)
    assert info == 0 and rinfo == 0
    assert rel(np.triu(out), np.triu(ref)) < TOL
    assert np.array_equal(np.tril(out, -1), junk)


def test_dpotrf_info(hip):
    a = spd_tile(64, 3)
    a[40, 40] = -1e6
    _, info = hip.dpotrf(a)
    assert info == 41


@pytest.mark.parametrize("n", TILE_SIZES)
def test_dgetrf_tile(hip, orc, n):
    a = rand_tile(n, 21, diag=n)
    out = hip.dgetrf(a)
    ref = orc.tile_getrf(a)
    assert rel(out, ref) < TOL


@pytest.mark.parametrize("n", [4, 32, 64, 96, 256, 512])
def test_qr_tiles(hip, orc, n):
    a = rand_tile(n, 31)
    A, T, tau = hip.dgeqrt(a)
    Ar, Tr, taur = orc.tile_geqrt(a)
    assert rel(A, Ar) < TOL and rel(np.triu(T), np.triu(Tr)) < TOL and rel(tau, taur) < TOL
    c = rand_tile(n, 32)
    assert rel(hip.dormqr(Ar, Tr, c), orc.tile_ormqr(Ar, Tr, c)) < TOL
    a2 = rand_tile(n, 33)
    r = np.triu(Ar)
    A1, A2, T2, tau2 = hip.dtsqrt(r, a2)
    A1r, A2r, T2r, tau2r = orc.tile_tsqrt(r, a2)
    assert rel(A1, A1r) < TOL and rel(A2, A2r) < TOL and rel(np.triu(T2), np.triu(T2r)) < TOL
    c1, c2 = rand_tile(n, 34), rand_tile(n, 35)
    o1, o2 = hip.dtsmqr(c1, c2, A2r, T2r)
    r1, r2 = orc.tile_tsmqr(c1, c2, A2r, T2r)
    assert rel(o1, r1) < TOL and rel(o2, r2) < TOL


@pytest.mark.parametrize("n", [64, 160])
def test_qr_tiles_zero_reflectors(hip, orc, n):
    """Columns whose Householder vector is zero give tau = 0 (dlarfg with a
    zero x), an identity reflector and a zero row and column of T."""
    a = np.triu(rand_tile(n, 41))
    a[:, 40:] = rand_tile(n, 42)[:, 40:]          # first 40 columns already upper triangular
    A, T, tau = hip.dgeqrt(a)
    Ar, Tr, taur = orc.tile_geqrt(a)
    assert np.all(taur[:40] == 0) and np.all(tau[:40] == 0)
    assert rel(A, Ar) < TOL and rel(np.triu(T), np.triu(Tr)) < TOL and rel(tau, taur) < TOL
    a2 = rand_tile(n, 43)
    a2[:, :10] = 0.0
    a2[:, 33] = 0.0                               # nonzero again after the first updates
    r = np.triu(Ar)
    A1, A2, T2, tau2 = hip.dtsqrt(r, a2)
    A1r, A2r, T2r, tau2r = orc.tile_tsqrt(r, a2)
    assert np.all(tau2r[:10] == 0) and np.all(tau2[:10] == 0)
    assert rel(A1, A1r) < TOL and rel(A2, A2r) < TOL and rel(np.triu(T2), np.triu(T2r)) < TOL
    assert rel(tau2, tau2r) < TOL
    assert np.array_equal(np.tril(A1, -1), np.tril(r, -1))   # below R: untouched


# ------------------------------------------------------------------ full algorithms
SIZES = [(4, 2), (256, 64), (512, 128), (1024, 256), (2048, 256), (2048, 512), (2048, 1024), (480, 96)]


@pytest.mark.parametrize("N,b", SIZES)
def test_cholesky(hip, orc, N, b):
    if N == 4:
        a = orc.gen("special4x4", 4, 2)
    else:
        a = orc.gen("sym", N, b, seed=0, shift=N)
    out = a.copy()
    info, rep = hip.cholesky(out, b=b)
    ref, rinfo = orc.cholesky(a, N, b)
    assert info == 0 and rinfo == 0
    A = orc.to_dense(a, N, b)
    U = orc.to_dense(out, N, b)
    Ur = orc.to_dense(ref, N, b)
    assert rel(np.triu(U), np.triu(Ur)) < TOL
    assert np.array_equal(np.tril(U, -1), np.tril(A, -1))     # lower part keeps its input
    assert orc.resid_chol(A, U) < 10


@pytest.mark.parametrize("N,b", SIZES)
def test_lu(hip, orc, N, b):
    a = orc.gen("dense", N, b, seed=0, shift=N)
    out = a.copy()
    info, rep, piv = hip.lu(out, b=b)
    ref = orc.lu(a, N, b)
    assert info == 0
    assert np.array_equal(piv, np.arange(1, N + 1))
    assert rel(out, ref) < TOL
    assert orc.resid_lu(orc.to_dense(a, N, b), orc.to_dense(out, N, b)) < 10


@pytest.mark.parametrize("N,b", [(4, 2), (256, 64), (512, 128), (1024, 256), (480, 96), (2048, 512)])
def test_qr(hip, orc, N, b):
    a = orc.gen("dense", N, b, seed=0, shift=0.0)
    out = a.copy()
    s = np.zeros(N * N)
    info, rep = hip.qr(out, s, b=b)
    ref, sref = orc.qr(a, N, b)
    assert rel(out, ref) < TOL
    MT = N // b
    for m in range(MT):
        for k in range(min(m, MT) + 1):
            off = b * b * (m + MT * k)
            if m == k or m > k:
                t = s[off:off + b * b].reshape(b, b).T   # column-major tile -> [r][c]
                tr = sref[off:off + b * b].reshape(b, b).T
                assert rel(np.triu(t), np.triu(tr)) < TOL, (m, k)
    r1, r2 = orc.resid_qr(orc.to_dense(a, N, b), out, s, N, b)     # ||A - QR||, ||I - Q^T Q|| (scaled)
    assert r1 < 10 and r2 < 10, (r1, r2)


@pytest.mark.parametrize("N,b", [(1024, 128), (2048, 256)])
def test_qr_schedules_bitwise(hip, orc, N, b):
    """The two-stream QR schedule (panel chain beside the updates), the
    in-order single-stream schedule and its graph replay give the same bits
    in A and in the T factors S."""
    a = orc.gen("dense", N, b, seed=5, shift=0.0)
    outs = []
    for kw in [{}, {"no_lookahead": True}, {"no_lookahead": True, "no_graph": True}, {}]:
        x, s = a.copy(), np.zeros(N * N)
        hip.qr(x, s, b=b, **kw)
        outs.append((x, s))
    for x, s in outs[1:]:
        assert np.array_equal(x.view(np.uint64), outs[0][0].view(np.uint64))
        assert np.array_equal(s.view(np.uint64), outs[0][1].view(np.uint64))
    ref, sref = orc.qr(a, N, b)
    assert rel(outs[0][0], ref) < TOL


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
def test_graph_and_streams_bitwise(hip, orc, algo):
    """Graph replay, stream-ordered launches and the no-look-ahead schedule give
    bitwise identical factors (kernels are deterministic, no atomics)."""
    N, b = 1024, 128
    a = orc.gen("sym" if algo == "cholesky" else "dense", N, b, seed=3, shift=N)
    outs = []
    for kw in [{}, {"no_graph": True}, {"no_lookahead": True}, {}]:
        x = a.copy()
        getattr(hip, algo)(x, b=b, **kw)
        outs.append(x)
    for x in outs[1:]:
        assert np.array_equal(x.view(np.uint64), outs[0].view(np.uint64))


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
def test_schedules_bitwise(hip, orc, algo):
    """The split-update chain schedule (default), the one-launch chain schedule,
    the row look-ahead schedule, with the bulk streams CU-masked (no graph) and
    without a reserve (captured graph), give bitwise identical FP64 factors.
    On the packed BF16 tier the solutions agree bitwise within each working
    storage: the FP32 working copy (split schedule, either reserve) and the
    FP64 one (both chain schedules)."""
    import os
    N, b = 1280, 256                       # MT = 5: every step kind, packed panels (b % 128 == 0)
    a = orc.gen("sym" if algo == "cholesky" else "dense", N, b, seed=5, shift=N)
    rhs = np.random.default_rng(2).standard_normal((N, 1))
    names = ("PALIAL_DAG", "PALIAL_RESERVE_CUS", "PALIAL_F32WORK")
    keep = {k: os.environ.get(k) for k in names}
    facs, sols = [], {}
    try:
        for dag, res, fw in [("2", "16", "1"), ("1", "16", "1"), ("0", "16", "1"), ("2", "0", "1"), ("1", "0", "1"),
                             ("2", "16", "0")]:
            hip.finalize()
            os.environ.update(dict(zip(names, (dag, res, fw))))
            hip.init(0)
            x = a.copy()
            getattr(hip, algo)(x, b=b)
            facs.append(x)
            if dag != "0":                 # the row look-ahead uses the unpacked BF16 kernel
                sol, rep, rc = hip.solve(a, rhs, algo=algo, precision=hip.PREC_BF16, b=b, factor_tile=b)
                assert rc == 0 and rep.backward_error <= 1e-14, (dag, res, fw)
                sols.setdefault(dag == "2" and fw == "1", []).append(sol)
    finally:
        hip.finalize()
        for k, v in keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        hip.init(0)
    for i in range(1, len(facs)):
        assert np.array_equal(facs[i].view(np.uint64), facs[0].view(np.uint64)), i
    assert len(sols[True]) == 2 and len(sols[False]) == 3
    for group in sols.values():
        for i in range(1, len(group)):
            assert np.array_equal(group[i].view(np.uint64), group[0].view(np.uint64)), i


def test_graph_cache_bounded(hip, orc):
    """More captured factorizations than the library keeps (16): the cache
    is emptied and refilled, and every replay still gives the oracle's
    factor (host buffers, so each call captures for its own shape)."""
    shapes = [(n * b, b) for b in (32, 64) for n in range(1, 11)]
    for N, b in shapes:
        a = orc.gen("sym", N, b, seed=N + b, shift=N)
        for _ in range(2):                      # capture, then replay
            x = a.copy()
            info, _ = hip.cholesky(x, b=b)
            ref, _ = orc.cholesky(a, N, b)
            assert info == 0 and rel(x, ref) < TOL, (N, b)
