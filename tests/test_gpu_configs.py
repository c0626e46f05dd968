"""GPU tests at the workload configurations' own sizes, through the C ABI.

  config 1  N=2048, b=256, seed 0: Cholesky (S + N I), LU (G + N I), QR (G)
            against the oracle's sequential program order
  config 2  N=8192, b=512: the three factorizations against the oracle's
            OpenMP DAG with OpenBLAS tile kernels; N=16384..32768: randomized
            residual probes computed on the host from copied-back tiles
  config 3  N=24576 and 32768, b=512, A = Q diag(s) Q^T (Q from the library's
            own QR of a seeded Gaussian, s geometric 1..1/kappa), kappa in
            {1e2, 1e5, 1e8}, 1 and 16 N(0,1) right-hand sides, Cholesky and
            LU solves in every precision tier; the backward error is
            recomputed on the host (numpy BLAS, no library kernel)
  config 4  the seeded 10,000-matrix batch (n ~ UniformInt[16,512], even
            Cholesky of S + n I, odd LU of G + n I), FP64 and FP32, every
            matrix checked

Bars: factors within 1e-10 relative of the oracle and scaled residuals
||A - F||_F / (n eps ||A||_F) below 10 (probes: ||(A - F) x||_2 with x ~
N(0, I), which estimates ||A - F||_F); solves: backward error
max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) <= 1e-14 as
the library reports it and <= 2e-14 recomputed on the host (different
summation order).
"""
import ctypes as C

import numpy as np
import pytest

from layout_probe import inf_norm, matvec

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
EPS32 = float(np.finfo(np.float32).eps)


def rel(x, ref):
    return np.abs(x - ref).max() / np.abs(ref).max()


def inputs(orc, algo, N, b, seed=0):
    if algo == "cholesky":
        return orc.gen("sym", N, b, seed=seed, shift=N)
    return orc.gen("dense", N, b, seed=seed, shift=N if algo == "lu" else 0.0)


def dense_residuals(hip, orc, algo, a, F, S, N, b):
    A = orc.to_dense(a, N, b)
    D = orc.to_dense(F, N, b)
    nA = np.linalg.norm(A)
    if algo == "cholesky":
        U = np.triu(D)
        return [np.linalg.norm(A - U.T @ U) / (N * EPS * nA)]
    if algo == "lu":
        L = np.tril(D, -1) + np.eye(N)
        return [np.linalg.norm(A - L @ np.triu(D)) / (N * EPS * nA)]
    Qt = orc.gen("identity", N, b)
    hip.qr_apply(F, S, Qt, trans=hip.NOTRANS, b=b, N=N)     # Q formed on the GPU
    Q = orc.to_dense(Qt, N, b)
    return [np.linalg.norm(A - Q @ np.triu(D)) / (N * EPS * nA),
            np.linalg.norm(np.eye(N) - Q.T @ Q) / (N * EPS)]


def factor_host(hip, algo, a, N, b):
    F = a.copy()
    S = np.zeros(N * N) if algo == "qr" else None
    if algo == "cholesky":
        info, _ = hip.cholesky(F, b=b)
    elif algo == "lu":
        info, _, piv = hip.lu(F, b=b)
        assert np.array_equal(piv, np.arange(1, N + 1))
    else:
        info, _ = hip.qr(F, S, b=b)
    assert info == 0
    return F, S


# ------------------------------------------------------------------ config 1
@pytest.mark.parametrize("algo", ["cholesky", "lu", "qr"])
def test_config1_vs_oracle(hip, orc, algo):
    N, b = 2048, 256
    a = inputs(orc, algo, N, b)
    F, S = factor_host(hip, algo, a, N, b)
    if algo == "cholesky":
        ref, info = orc.cholesky(a, N, b)
        assert info == 0
    elif algo == "lu":
        ref = orc.lu(a, N, b)
    else:
        ref, Sref = orc.qr(a, N, b)
    assert rel(F, ref) < 1e-10
    for r in dense_residuals(hip, orc, algo, a, F, S, N, b):
        assert r < 10


# ------------------------------------------------------------------ config 2, N = 8192
@pytest.fixture
def openblas(orc):
    assert orc.use_openblas(), "the oracle's OpenBLAS back end (scipy's libscipy_openblas) is missing"
    orc.openblas_threads(1)
    yield orc
    orc.use_naive()


@pytest.mark.parametrize("algo", ["cholesky", "lu", "qr"])
def test_config2_n8192_vs_openblas_oracle(hip, openblas, algo):
    orc = openblas
    N, b = 8192, 512
    a = inputs(orc, algo, N, b)
    F, S = factor_host(hip, algo, a, N, b)
    if algo == "cholesky":
        ref, _ = orc.cholesky(a, N, b, omp_threads=16)
    elif algo == "lu":
        ref = orc.lu(a, N, b, omp_threads=16)
    else:
        ref, _ = orc.qr(a, N, b, omp_threads=16)
    err = rel(F, ref)
    res = dense_residuals(hip, orc, algo, a, F, S, N, b)
    print(f"config2 {algo} N={N}: rel {err:.2e}, scaled residuals {res}", flush=True)
    assert err < 1e-10
    for r in res:
        assert r < 10


# ------------------------------------------------------------------ config 2, N >= 16384
def to_tile_col(X, N, b):
    """N x k (k <= b) -> N x b layout-L tile column (extra columns zero)."""
    MT, k = N // b, X.shape[1]
    T = np.zeros((MT, b, b))                 # [m, c, r]
    T[:, :k, :] = X.reshape(MT, b, k).transpose(0, 2, 1)
    return np.ascontiguousarray(T.reshape(-1))


def from_tile_col(t, N, b, k):
    MT = N // b
    return t.reshape(MT, b, b)[:, :k, :].transpose(0, 2, 1).reshape(N, k)


@pytest.mark.parametrize("N", [16384, 24576, 32768])
@pytest.mark.parametrize("algo", ["cholesky", "lu", "qr"])
def test_config2_large_probes(hip, orc, algo, N):
    b = 512
    A = hip.DeviceMatrix(N, b)
    S = hip.DeviceMatrix(N, b) if algo == "qr" else None
    dC = C.c_void_p()
    try:
        A.generate(hip.GEN_SYM if algo == "cholesky" else hip.GEN_DENSE, seed=0,
                   shift=0.0 if algo == "qr" else float(N))
        a = A.to_host()
        info, rep = hip.factor_device(algo, A, S=S)
        assert info == 0
        f = A.to_host()
        X = np.random.default_rng(7).standard_normal((N, 2))
        Y = matvec(a, N, b, X)
        nA = np.linalg.norm(a)
        res = []
        if algo == "cholesky":
            Z = matvec(f, N, b, matvec(f, N, b, X, "upper"), "upper_t")
            res.append(np.linalg.norm(Y - Z, axis=0).max() / (N * EPS * nA))
        elif algo == "lu":
            Z = matvec(f, N, b, matvec(f, N, b, X, "upper"), "unit_lower")
            res.append(np.linalg.norm(Y - Z, axis=0).max() / (N * EPS * nA))
        else:
            # (A - QR) x = A x - Q (R x) and (I - Q^T Q) x, Q applied on the GPU
            RX = matvec(f, N, b, X, "upper")
            hip._check(hip.lib().palial_hip_malloc(C.byref(dC), N * b * 8))
            t = to_tile_col(RX, N, b)
            hip._check(hip.lib().palial_hip_memcpy_h2d(dC, t.ctypes.data, t.nbytes))
            hip.qr_apply(A, S, dC.value, trans=hip.NOTRANS, ncols=b)
            hip._check(hip.lib().palial_hip_memcpy_d2h(t.ctypes.data, dC, t.nbytes))
            QRX = from_tile_col(t, N, b, 2)
            res.append(np.linalg.norm(Y - QRX, axis=0).max() / (N * EPS * nA))
            t = to_tile_col(X, N, b)
            hip._check(hip.lib().palial_hip_memcpy_h2d(dC, t.ctypes.data, t.nbytes))
            hip.qr_apply(A, S, dC.value, trans=hip.NOTRANS, ncols=b)
            hip.qr_apply(A, S, dC.value, trans=hip.TRANS, ncols=b)
            hip._check(hip.lib().palial_hip_memcpy_d2h(t.ctypes.data, dC, t.nbytes))
            QtQX = from_tile_col(t, N, b, 2)
            res.append(np.linalg.norm(QtQX - X, axis=0).max() / (N * EPS))
        print(f"config2 {algo} N={N}: factor {rep.factor_ms:.1f} ms, probe residuals {res}", flush=True)
        for r in res:
            assert r < 10
    finally:
        if dC.value:
            hip.lib().palial_hip_free(dC)
        A.free()
        if S is not None:
            S.free()


# ------------------------------------------------------------------ config 3
TIERS = {"fp64": 0, "fp32": 1, "bf16": 2, "fp8": 3, "bf16x3": 4}


@pytest.fixture(scope="module")
def cond_gens(hip):
    gens = {}

    def get(N):
        if N not in gens:
            for g in gens.values():
                g.free()
            gens.clear()
            gens[N] = hip.CondGenerator(N, 512, seed=0)
        return gens[N]
    yield get
    for g in gens.values():
        g.free()


@pytest.mark.parametrize("kappa", [1e2, 1e5, 1e8])
@pytest.mark.parametrize("N", [24576, 32768])
def test_config3_solves(hip, cond_gens, kappa, N):
    b = 512
    gen = cond_gens(N)
    A = hip.DeviceMatrix(N, b)
    dB = C.c_void_p()
    L = hip.lib()
    try:
        gen.fill(A, kappa)
        a = A.to_host()
        nA = inf_norm(a, N, b)
        hip._check(L.palial_hip_malloc(C.byref(dB), N * 16 * 8))
        for nrhs in (1, 16):
            Bh = np.asfortranarray(np.random.default_rng(1).standard_normal((N, nrhs)))
            for algo in ("cholesky", "lu"):
                sols, reps = {}, {}
                for name, prec in TIERS.items():
                    hip._check(L.palial_hip_memcpy_h2d(dB, Bh.ctypes.data, Bh.nbytes))
                    _, rep, rc = hip.solve(A, (dB, nrhs, N), algo=algo, precision=prec, device=True, max_iter=60,
                                           check=False, factor_tile=b)
                    assert rc == 0, (algo, name, rc, rep.backward_error)
                    X = np.empty((N, nrhs), order="F")
                    hip._check(L.palial_hip_memcpy_d2h(X.ctypes.data, dB, X.nbytes))
                    sols[name], reps[name] = X, rep
                # one host pass over A for every tier's solution
                Xall = np.hstack([sols[n] for n in TIERS])
                R = np.tile(Bh, len(TIERS)) - matvec(a, N, b, Xall)
                for i, name in enumerate(TIERS):
                    sl = slice(i * nrhs, (i + 1) * nrhs)
                    X, Rn = Xall[:, sl], R[:, sl]
                    eta = (np.abs(Rn).max(0) / (nA * np.abs(X).max(0) + np.abs(Bh).max(0))).max()
                    rep = reps[name]
                    print(f"config3 N={N} kappa={kappa:g} nrhs={nrhs} {algo} {name}: "
                          f"eta lib {rep.backward_error:.2e} host {eta:.2e}, factor {rep.factor_ms:.1f} ms, "
                          f"refine {rep.solve_ms:.1f} ms, {rep.iterations} it + {rep.gmres_iterations} GMRES, "
                          f"tier used {rep.precision_used}, shifts {rep.shifts}", flush=True)
                    assert rep.backward_error <= 1e-14
                    assert eta <= 2e-14
    finally:
        if dB.value:
            L.palial_hip_free(dB)
        A.free()


# ------------------------------------------------------------------ config 4
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_config4_seeded_batch(hip, orc, prec):
    count = 10000
    rng = np.random.default_rng(0)
    n = rng.integers(16, 513, count).astype(np.int32)
    kind = (np.arange(count) % 2).astype(np.int32)
    rhs = rng.standard_normal(int(n.sum()))
    p = hip.PREC_FP64 if prec == "fp64" else hip.PREC_FP32
    bt = hip.Batch(n, kind, precision=p)
    try:
        bt.generate(0)
        bt.set(rhs=[rhs[o:o + m] for o, m in zip(bt.boff, n)])
        A0, B0 = bt.get()
        rep = bt.factor(solve=True)
        F, X = bt.get()
        info = bt.info.copy()
    finally:
        bt.free()
    assert not info.any()
    eps = EPS if prec == "fp64" else EPS32
    worst_f = worst_x = 0.0
    probe = np.random.default_rng(3)
    for i in range(count):
        m = int(n[i])
        A = A0[i].astype(np.float64)
        if i % 97 == 0:    # the batch generator against the oracle's counter-based uniform
            r, c = (i * 7) % m, (i * 13) % m
            u = orc.uniform(i, min(r, c), max(r, c)) if kind[i] == 0 else orc.uniform(i, r, c)
            assert A0[i][r, c] == np.asarray(u + (m if r == c else 0.0), dtype=A0[i].dtype)
        Fi = F[i].astype(np.float64)
        v = probe.standard_normal(m)
        if kind[i] == 0:
            U = np.triu(Fi)
            d = A @ v - U.T @ (U @ v)
            assert np.array_equal(np.tril(F[i], -1), np.tril(A0[i], -1))   # strict lower part untouched
        else:
            Lw = np.tril(Fi, -1) + np.eye(m)
            d = A @ v - Lw @ (np.triu(Fi) @ v)
        worst_f = max(worst_f, np.linalg.norm(d) / (m * eps * np.linalg.norm(A)))
        x = X[i].astype(np.float64)
        b = B0[i].astype(np.float64)
        eta = np.abs(b - A @ x).max() / (np.abs(A).sum(1).max() * np.abs(x).max() + np.abs(b).max())
        worst_x = max(worst_x, eta / eps)
    print(f"config4 {prec}: {rep.factor_ms:.2f} ms, worst scaled factor probe {worst_f:.2f}, "
          f"worst backward error {worst_x:.2f} eps", flush=True)
    assert worst_f < 10
    assert worst_x < 10 * 512


# ------------------------------------------------------------------ GETRF pivot rules on the GPU
def blockdiag_case(n, pivot):
    """Block-diagonal tile whose second 32x32 block starts with the given
    pivot: the first block's elimination adds exact zeros to the second, so
    its leading pivot reaches GETRF exactly as written."""
    r = np.random.default_rng(6)
    A = np.zeros((n, n))
    A[:32, :32] = r.uniform(-0.5, 0.5, (32, 32)) + 32 * np.eye(32)
    A[32:, 32:] = r.uniform(-0.5, 0.5, (n - 32, n - 32)) + n * np.eye(n - 32)
    if pivot == "zero":
        A[32, 32] = 0.0
    else:      # subnormal pivot and subnormal column: divide finite, 1/pivot overflows
        A[32, 32] = 2.0 ** -1030
        A[33:, 32] = np.arange(1, n - 32) * 2.0 ** -1030
        # a small pivot row keeps the rank-1 update by multipliers up to n
        # from destroying the diagonal dominance of the rest (no growth)
        A[32, 33:] *= 2.0 ** -12
    return np.asfortranarray(A)


@pytest.mark.parametrize("n", [64, 128, 512])
@pytest.mark.parametrize("pivot", ["zero", "subnormal"])
def test_getrf_pivot_rules_on_gpu(hip, orc, n, pivot):
    """src/srclu/dgetrf_no_pivot.h:21-39: a zero pivot is skipped (no
    scaling), a pivot at or below the safe minimum divides instead of
    multiplying by its reciprocal."""
    A = blockdiag_case(n, pivot)
    got = hip.dgetrf(A)
    ref = orc.tile_getrf(A)
    assert np.array_equal(np.isfinite(got), np.isfinite(ref))
    assert np.isfinite(ref[33:, 32]).all()
    fin = np.isfinite(ref)
    assert np.array_equal(got[fin], ref[fin]) or np.abs(got[fin] - ref[fin]).max() <= 1e-12 * np.abs(ref[fin]).max()
    if pivot == "subnormal":
        assert np.array_equal(got[33:, 32], np.arange(1, n - 32, dtype=np.float64))
    else:
        assert np.array_equal(got[33:, 32], A[33:, 32])


# ------------------------------------------------------------------ the drop-in CLI
@pytest.mark.parametrize("algo", ["cholesky", "lu", "qr"])
def test_cli_runs_on_gpu(hip, algo):
    import subprocess
    exe = hip.DRIVER_PATH
    r = subprocess.run([exe, "-a", algo, "-m", "1024", "-b", "256", "-v"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    if algo == "qr":
        assert r.stdout == ""                                      # silent, as main.c:161-167
    else:
        eta = float(r.stdout.split("backward error ")[1].split()[0])
        assert eta <= 1e-14
    r = subprocess.run([exe, "-a", algo, "-m", "1024", "-b", "256", "-t"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.startswith(f"{algo}, 1024, 256, ")
