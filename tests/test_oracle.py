"""CPU tests: pin the oracle before trusting it.

Each oracle function is checked against something that does not restate its
own formula: LAPACK (through scipy) on the same inputs, closed forms, the
reference's own fixtures (exp/data task order, SPECIAL4x4, IDENTITY) and
mathematical invariants (uniqueness of Cholesky / no-pivot LU, R^T R = A^T A).
"""
import hashlib
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def rel(x, ref):
    return np.abs(np.asarray(x) - ref).max() / max(np.abs(ref).max(), 1e-300)


@pytest.fixture(params=["naive", "openblas"])
def backend(request, orc):
    if request.param == "openblas":
        if not orc.use_openblas():
            pytest.skip("no LP64 OpenBLAS found")
    else:
        orc.use_naive()
    yield request.param
    orc.use_naive()


# ------------------------------------------------------------------ layout
def test_tile_address_closed_form(orc):
    # include/tile_address.h:15 : offset = b*b*(m + (N/b)*n)
    for N, b in [(4, 2), (12, 4), (64, 16)]:
        MT = N // b
        seen = set()
        for n in range(MT):
            for m in range(MT):
                off = orc.lib().orc_tile_offset(N, b, m, n)
                assert off == b * b * (m + MT * n)
                seen.add(off)
        assert sorted(seen) == [b * b * t for t in range(MT * MT)]   # tiles tile the buffer


def test_layout_matches_print_map(orc):
    # include/print.h:27 maps tile (i,j) element dA[k*b+l] to B[k*N + l + b*j*N + i*b]
    N, b = 12, 4
    t = np.arange(N * N, dtype=np.float64)
    MT = N // b
    B = np.zeros(N * N)
    for i in range(MT):
        for j in range(MT):
            base = b * b * (i + MT * j)
            for k in range(b):
                for l in range(b):
                    B[k * N + l + b * j * N + i * b] = t[base + k * b + l]
    dense = orc.to_dense(t, N, b)            # dense[row, col]
    assert np.array_equal(B.reshape(N, N).T, dense)   # B is column-major
    assert np.array_equal(orc.from_dense(dense, N, b), t)


# ------------------------------------------------------------------ generators
def test_generators(orc):
    N, b = 64, 16
    S = orc.to_dense(orc.gen("sym", N, b, seed=3, shift=0.0), N, b)
    assert np.array_equal(S, S.T)
    assert S.min() >= -0.5 and S.max() < 0.5
    G = orc.to_dense(orc.gen("dense", N, b, seed=3, shift=5.0), N, b)
    assert np.allclose(np.diag(G) - 5.0, [orc.uniform(3, i, i) for i in range(N)], rtol=0, atol=1e-15)
    # layout independence: the same seeded matrix for any tile size
    G2 = orc.to_dense(orc.gen("dense", N, 8, seed=3, shift=5.0), N, 8)
    assert np.array_equal(G, G2)
    P = orc.to_dense(orc.gen("palial", N, b, seed=1), N, b)
    off = P[~np.eye(N, dtype=bool)]
    assert np.all(off == 0.5) and np.all((np.diag(P) >= 173) & (np.diag(P) <= 346))
    assert np.array_equal(orc.to_dense(orc.gen("identity", N, b), N, b), np.eye(N))


def test_uniform_known_values(orc):
    # regression pin of the counter-based generator (splitmix64)
    def sm64(z):
        M = (1 << 64) - 1
        z = (z + 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    for seed, i, j in [(0, 0, 0), (0, 5, 7), (123, 1000, 3), (7, 32767, 32767)]:
        x = sm64((sm64(seed) + ((i << 32) | j)) & ((1 << 64) - 1))
        assert orc.uniform(seed, i, j) == (x >> 11) * 2.0 ** -53 - 0.5


# ------------------------------------------------------------------ task order (reference traces)
def test_cholesky_task_order_matches_reference_traces(orc):
    g = json.load(open(os.path.join(GOLD, "chol_task_names.json")))
    assert orc.task_names("cholesky", 16) == g["names_MT16"]
    for MT, h in g["sha256"].items():
        names = orc.task_names("cholesky", int(MT))
        assert len(names) == int(MT) + int(MT) * (int(MT) - 1) + int(MT) * (int(MT) - 1) * (int(MT) - 2) // 6
        assert hashlib.sha256("\n".join(names).encode()).hexdigest() == h


def test_lu_qr_task_counts(orc):
    for MT in (1, 2, 5, 16):
        n_lu = orc.task_names("lu", MT)
        n_qr = orc.task_names("qr", MT)
        expect = MT + MT * (MT - 1) + (MT - 1) * MT * (2 * MT - 1) // 6
        assert len(n_lu) == expect and len(n_qr) == expect
        assert len(set(n_qr)) == len(n_qr) or MT > 9   # names are not unique once indices reach 10


# ------------------------------------------------------------------ tile kernels vs LAPACK
GT = np.load(os.path.join(GOLD, "lapack_tiles.npz"))


@pytest.mark.parametrize("n", [4, 8, 32])
def test_tiles_vs_lapack(orc, backend, n):
    U, info = orc.tile_potrf(GT[f"potrf_in_{n}"])
    assert info == 0 and rel(np.triu(U), GT[f"potrf_out_{n}"]) < 1e-13
    A, T, tau = orc.tile_geqrt(GT[f"geqrt_in_{n}"])
    assert rel(A, GT[f"geqrt_A_{n}"]) < 1e-13 and rel(np.triu(T), GT[f"geqrt_T_{n}"]) < 1e-13
    A1, A2, T2, tau2 = orc.tile_tsqrt(GT[f"tsqrt_R_{n}"], GT[f"tsqrt_A2_{n}"])
    assert rel(A1, GT[f"tsqrt_A1out_{n}"]) < 1e-13 and rel(A2, GT[f"tsqrt_A2out_{n}"]) < 1e-13
    assert rel(np.triu(T2), GT[f"tsqrt_T_{n}"]) < 1e-13
    o1, o2 = orc.tile_tsmqr(GT[f"tsmqr_C1_{n}"], GT[f"tsmqr_C2_{n}"], GT[f"tsqrt_A2out_{n}"], GT[f"tsqrt_T_{n}"])
    assert rel(o1, GT[f"tsmqr_out1_{n}"]) < 1e-13 and rel(o2, GT[f"tsmqr_out2_{n}"]) < 1e-13
    c = orc.tile_ormqr(GT[f"geqrt_A_{n}"], GT[f"geqrt_T_{n}"], GT[f"ormqr_C_{n}"])
    assert rel(c, GT[f"ormqr_out_{n}"]) < 1e-13


@pytest.mark.parametrize("n", [3, 16, 48])
def test_blas_tasks_vs_numpy(orc, backend, n):
    r = np.random.default_rng(n)
    A = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    B = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    Cm = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    U = np.triu(A) + 2 * np.eye(n)
    Ljunk = np.tril(A, -1) + np.triu(B)                  # strict lower = L, rest junk
    assert rel(orc.chol_trsm(U + np.tril(B, -1), Cm), sl.solve_triangular(U, Cm, trans="T")) < 1e-12
    syrk = orc.chol_syrk(A, Cm)
    assert rel(np.triu(syrk), np.triu(Cm - A.T @ A)) < 1e-13 and np.array_equal(np.tril(syrk, -1), np.tril(Cm, -1))
    assert rel(orc.chol_gemm(A, B, Cm), Cm - A.T @ B) < 1e-13
    assert rel(orc.lu_trsm_r(U + np.tril(B, -1), Cm), Cm @ np.linalg.inv(U)) < 1e-11
    L = np.tril(A, -1) + np.eye(n)
    assert rel(orc.lu_trsm_l(Ljunk, Cm), np.linalg.solve(L, Cm)) < 1e-11
    assert rel(orc.lu_gemm(A, B, Cm), Cm - A @ B) < 1e-13


def test_getrf_no_pivot(orc, backend):
    # src/srclu/dgetrf_no_pivot.h: on a diagonally dominant tile GEPP picks the diagonal
    n = 32
    r = np.random.default_rng(5)
    a = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)) + n * np.eye(n))
    f = orc.tile_getrf(a)
    P, L, U = sl.lu(a)
    assert np.array_equal(P, np.eye(n))
    assert rel(np.triu(f), U) < 1e-13 and rel(np.tril(f, -1), np.tril(L, -1)) < 1e-13


def test_getrf_zero_pivot_skipped(orc):
    # alpha == 0 is skipped (no scaling, no error), dgetrf_no_pivot.h:21
    a = np.asfortranarray(np.array([[0.0, 1.0], [2.0, 3.0]]))
    f = orc.tile_getrf(a)
    assert f[1, 0] == 2.0 and f[1, 1] == 3.0 - 2.0 * 1.0


# ------------------------------------------------------------------ full algorithms
@pytest.mark.parametrize("N,b", [(64, 16), (192, 64)])
def test_cholesky_unique_factor(orc, backend, N, b):
    a = orc.gen("sym", N, b, seed=0, shift=N)
    u, info = orc.cholesky(a, N, b)
    A = orc.to_dense(a, N, b)
    U = orc.to_dense(u, N, b)
    assert info == 0
    assert rel(np.triu(U), sl.cholesky(A)) < 1e-13
    assert np.array_equal(np.tril(U, -1), np.tril(A, -1))
    assert orc.resid_chol(A, U) < 10


def test_cholesky_closed_form_populate(orc):
    """A = D + 1/2 * 1 1^T (include/populate.h generator) has an O(N) Cholesky:
    gamma_1 = 1, dt_j = d_j + gamma_j/2, U_jj = sqrt(dt_j), U_ji = gamma_j/(2 sqrt(dt_j)) (i > j),
    gamma_{j+1} = gamma_j d_j / dt_j."""
    N, b = 256, 32
    a = orc.gen("palial", N, b, seed=9)
    A = orc.to_dense(a, N, b)
    d = np.diag(A) - 0.5
    u, _ = orc.cholesky(a, N, b)
    U = np.triu(orc.to_dense(u, N, b))
    g = 1.0
    for j in range(N):
        dt = d[j] + 0.5 * g
        assert abs(U[j, j] - np.sqrt(dt)) < 1e-13 * np.sqrt(dt)
        if j + 1 < N:
            assert np.allclose(U[j, j + 1:], 0.5 * g / np.sqrt(dt), rtol=1e-12, atol=0)
        g = g * d[j] / dt


def test_special4x4_and_identity(orc):
    a = orc.gen("special4x4", 4, 2)
    A = orc.to_dense(a, 4, 2)
    u, info = orc.cholesky(a, 4, 2)
    assert info == 0 and rel(np.triu(orc.to_dense(u, 4, 2)), sl.cholesky(A)) < 1e-14
    N, b = 16, 4
    I = orc.gen("identity", N, b)
    u, _ = orc.cholesky(I, N, b)
    assert np.array_equal(u, I)
    assert np.array_equal(orc.lu(I, N, b), I)
    r, s = orc.qr(I, N, b)
    assert np.array_equal(r, I) and not s.any()            # R = I, V = 0, tau = 0, T = 0


@pytest.mark.parametrize("N,b", [(64, 16), (192, 64)])
def test_lu_unique_factor(orc, backend, N, b):
    a = orc.gen("dense", N, b, seed=0, shift=N)
    f = orc.to_dense(orc.lu(a, N, b), N, b)
    A = orc.to_dense(a, N, b)
    P, L, U = sl.lu(A)
    assert np.array_equal(P, np.eye(N))
    assert rel(np.triu(f), U) < 1e-12 and rel(np.tril(f, -1), np.tril(L, -1)) < 1e-12
    assert orc.resid_lu(A, f) < 10


@pytest.mark.parametrize("N,b", [(64, 16), (128, 32)])
def test_qr_invariants(orc, backend, N, b):
    a = orc.gen("dense", N, b, seed=0, shift=0.0)
    r, s = orc.qr(a, N, b)
    A = orc.to_dense(a, N, b)
    R = np.triu(orc.to_dense(r, N, b))
    Rref = sl.qr(A, mode="r")[0]
    sg = np.sign(np.diag(R)) * np.sign(np.diag(Rref))
    assert rel(R, sg[:, None] * Rref) < 1e-12
    assert rel(R.T @ R, A.T @ A) < 1e-13
    r1, r2 = orc.resid_qr(A, r, s, N, b)            # Q formed from the reflectors and T factors
    assert r1 < 10 and r2 < 10
    bad = s.copy()
    bad[b * b * 1:b * b * 1 + b * b] *= 1.01       # a perturbed T factor (tile (1, 0)) breaks Q
    assert max(orc.resid_qr(A, r, bad, N, b)) > 1e3


def test_backends_agree(orc):
    if not orc.use_openblas():
        pytest.skip("no OpenBLAS")
    N, b = 128, 32
    outs = []
    for be in ("openblas", "naive"):
        orc.use_openblas() if be == "openblas" else orc.use_naive()
        outs.append((orc.cholesky(orc.gen("sym", N, b, 1, N), N, b)[0], orc.lu(orc.gen("dense", N, b, 1, N), N, b),
                     orc.qr(orc.gen("dense", N, b, 1, 0.0), N, b)))
    orc.use_naive()
    (c1, l1, (q1, s1)), (c2, l2, (q2, s2)) = outs
    assert rel(c1, c2) < 1e-13 and rel(l1, l2) < 1e-13 and rel(q1, q2) < 1e-12 and rel(s1, s2) < 1e-12


def test_omp_dag_equals_sequential(orc):
    N, b = 128, 32
    for algo in ("cholesky", "lu", "qr"):
        kind = "sym" if algo == "cholesky" else "dense"
        a = orc.gen(kind, N, b, 2, N if algo != "qr" else 0.0)
        f = getattr(orc, algo)
        seq = f(a, N, b)
        par = f(a, N, b, omp_threads=4)
        seq = seq[0] if isinstance(seq, tuple) else seq
        par = par[0] if isinstance(par, tuple) else par
        assert np.array_equal(seq, par)   # each tile sees the same operation sequence


def test_solves(orc):
    N, b = 96, 32
    a = orc.gen("sym", N, b, 4, N)
    A = orc.to_dense(a, N, b)
    rhs = np.random.default_rng(0).standard_normal((N, 2))
    u, _ = orc.cholesky(a, N, b)
    x = orc.chol_solve(u, N, b, rhs)
    assert rel(x, np.linalg.solve(A, rhs)) < 1e-13
    g = orc.gen("dense", N, b, 4, N)
    G = orc.to_dense(g, N, b)
    x = orc.lu_solve(orc.lu(g, N, b), N, b, rhs)
    assert rel(x, np.linalg.solve(G, rhs)) < 1e-13


# ---------------------------------------------------------------- Q application, config-3 inputs
@pytest.mark.parametrize("N,b", [(64, 16), (96, 32)])
def test_apply_q_invariants(orc, N, b):
    """The oracle's Q application is pinned by the factorization it inverts:
    Q^T A = R and Q R = A (a reversed order, a dropped block or T in place of
    T^T breaks both), Q orthogonal, and Q^T (Q C) = C."""
    a = orc.gen("dense", N, b, seed=3, shift=0.0)
    F, S = orc.qr(a, N, b)
    A = orc.to_dense(a, N, b)
    R = np.triu(orc.to_dense(F, N, b))
    assert np.abs(orc.apply_q(F, S, N, b, A, trans=True) - R).max() < 1e-13 * N
    assert np.abs(orc.apply_q(F, S, N, b, R, trans=False) - A).max() < 1e-13 * N
    Q = orc.form_q(F, S, N, b)
    assert np.abs(Q.T @ Q - np.eye(N)).max() < 1e-14 * N
    Cm = np.random.default_rng(0).standard_normal((N, 2 * b))
    back = orc.apply_q(F, S, N, b, orc.apply_q(F, S, N, b, Cm, trans=False), trans=True)
    assert np.abs(back - Cm).max() < 1e-13 * N


def test_gauss_generator_is_standard_normal(orc):
    """N(0,1) by Box-Muller: moments and a Kolmogorov-Smirnov test on 2^18
    entries, plus the closed form of one entry recomputed from its counter."""
    import math
    from scipy import stats
    N, b = 512, 128
    g = orc.gen("gauss", N, b, seed=11)
    n = g.size
    assert abs(g.mean()) < 5 / math.sqrt(n)
    assert abs(g.var() - 1.0) < 5 * math.sqrt(2.0 / n)
    assert stats.kstest(g, "norm").pvalue > 1e-3

    def sm64(z):
        M = (1 << 64) - 1
        z = (z + 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    i, j = 5, 300
    h = sm64((sm64(11) + ((i << 32) | j)) & ((1 << 64) - 1))
    u1 = ((h >> 11) + 1) * 2.0 ** -53
    u2 = (sm64(h) >> 11) * 2.0 ** -53
    z = math.sqrt(-2 * math.log(u1)) * math.cos(2 * math.pi * u2)
    D = orc.to_dense(g, N, b)
    assert D[i, j] == pytest.approx(z, rel=1e-15, abs=1e-15)
    assert orc.gauss(11, i, j) == D[i, j]


def test_geom_diag_spectrum(orc):
    N, b, kappa = 256, 64, 1e8
    d = np.diag(orc.to_dense(orc.gen("geomdiag", N, b, shift=kappa), N, b))
    assert d[0] == 1.0 and d[-1] == pytest.approx(1 / kappa, rel=4e-16)
    r = d[1:] / d[:-1]
    assert np.allclose(r, kappa ** (-1.0 / (N - 1)), rtol=1e-13)
    off = orc.to_dense(orc.gen("geomdiag", N, b, shift=kappa), N, b) - np.diag(d)
    assert not off.any()
