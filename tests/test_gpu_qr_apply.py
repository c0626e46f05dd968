"""GPU tests of the Q application (Left, Trans and NoTrans), Q formation, the
QR solve, the in-place transpose and the config-3 input generator, through
the C ABI against the CPU oracle.

Bars: Q^T C and Q C within 1e-12 (max-abs over max-abs) of the oracle
applying the same reflectors; the GPU-formed Q within 1e-10 of the oracle's Q
(the oracle factoring the same seeded input); scaled residuals
||A - QR||_F/(n eps ||A||_F) and ||I - Q^T Q||_F/(n eps) below 10; the QR
solve within 1e-10 relative of numpy's least-squares solution and a backward
error below 1e-14 x n; transposes bit-exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def rel(x, ref):
    return np.abs(x - ref).max() / np.abs(ref).max()


def gpu_qr(hip, orc, N, b, seed=0, kind="dense"):
    a = orc.gen(kind, N, b, seed=seed, shift=0.0)
    F = a.copy()
    S = np.zeros(N * N)
    hip.qr(F, S, b=b)
    return a, F, S


@pytest.mark.parametrize("N,b,ncols", [(512, 128, 512), (768, 256, 256), (1024, 256, 2048), (384, 96, 96),
                                       (512, 64, 64)])
@pytest.mark.parametrize("trans", ["T", "N"])
def test_qr_apply_vs_oracle(hip, orc, N, b, ncols, trans):
    _, F, S = gpu_qr(hip, orc, N, b, seed=1)
    Cd = np.random.default_rng(2).standard_normal((N, ncols))
    # N x ncols dense -> layout L with N/b row tiles and ncols/b column tiles
    MT, NC = N // b, ncols // b
    Ct = np.ascontiguousarray(Cd.reshape(MT, b, NC, b).transpose(2, 0, 3, 1).reshape(-1))
    tr = hip.TRANS if trans == "T" else hip.NOTRANS
    hip.qr_apply(F, S, Ct, trans=tr, b=b, N=N, ncols=ncols)
    got = Ct.reshape(NC, MT, b, b).transpose(1, 3, 0, 2).reshape(N, ncols)
    ref = orc.apply_q(F, S, N, b, Cd, trans=(trans == "T"))
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("N,b", [(2048, 256), (1024, 128)])
def test_form_q_vs_oracle_and_residuals(hip, orc, N, b):
    a, F, S = gpu_qr(hip, orc, N, b, seed=0)
    Qt = orc.gen("identity", N, b)
    hip.qr_apply(F, S, Qt, trans=hip.NOTRANS, b=b, N=N)
    Q = orc.to_dense(Qt, N, b)
    Fo, So = orc.qr(a, N, b)
    assert rel(Q, orc.form_q(Fo, So, N, b)) < 1e-10
    A = orc.to_dense(a, N, b)
    R = np.triu(orc.to_dense(F, N, b))
    assert np.linalg.norm(A - Q @ R) / (N * EPS * np.linalg.norm(A)) < 10
    assert np.linalg.norm(np.eye(N) - Q.T @ Q) / (N * EPS) < 10


@pytest.mark.parametrize("N,b,nrhs", [(1024, 128, 1), (1024, 256, 16), (1536, 512, 3), (512, 64, 5)])
def test_qr_solve_vs_lstsq(hip, orc, N, b, nrhs):
    a, F, S = gpu_qr(hip, orc, N, b, seed=4)
    F0, S0 = F.copy(), S.copy()
    A = orc.to_dense(a, N, b)
    B = np.random.default_rng(5).standard_normal((N, nrhs))
    X, rep = hip.qr_solve(F, S, B, b=b, N=N)
    ref = np.linalg.lstsq(A, B, rcond=None)[0]
    assert rel(X, ref) < 1e-10
    eta = np.abs(B - A @ X).max() / (np.abs(A).sum(1).max() * np.abs(X).max() + np.abs(B).max())
    assert eta < 1e-14 * N
    assert np.array_equal(F, F0) and np.array_equal(S, S0)     # the factors are inputs only


@pytest.mark.parametrize("N,b", [(512, 128), (768, 96), (256, 256)])
def test_transpose_and_symmetrize(hip, orc, N, b):
    g = orc.gen("dense", N, b, seed=9, shift=0.0)
    G = orc.to_dense(g, N, b)
    t = g.copy()
    hip.transpose(t, b=b, N=N)
    assert np.array_equal(orc.to_dense(t, N, b), G.T)
    s = g.copy()
    hip.transpose(s, average=True, b=b, N=N)
    Sd = orc.to_dense(s, N, b)
    assert np.array_equal(Sd, Sd.T)
    assert np.array_equal(Sd, 0.5 * (G + G.T))


def test_gauss_and_spectrum_generators(hip, orc):
    N, b = 1024, 256
    g = np.zeros(N * N)
    hip.generate(g, hip.GEN_GAUSS, seed=11, N=N, b=b)
    ref = orc.gen("gauss", N, b, seed=11)
    # device and host libm agree to a few ulp of log / cos
    assert np.abs(g - ref).max() < 1e-13
    d = np.zeros(N * N)
    hip.generate(d, hip.GEN_GEOM_DIAG, shift=1e5, N=N, b=b)
    dref = orc.gen("geomdiag", N, b, shift=1e5)
    assert np.abs(d - dref).max() <= 4 * EPS


@pytest.mark.parametrize("kappa", [1e2, 1e8])
def test_cond_generator_spectrum(hip, orc, kappa):
    """A = Q diag(s) Q^T from the library's own QR of a Gaussian: symmetric,
    eigenvalues s (geometric from 1 to 1/kappa) to within n eps, Q orthogonal."""
    N, b = 1024, 256
    gen = hip.CondGenerator(N, b, seed=3)
    A = hip.DeviceMatrix(N, b)
    try:
        gen.fill(A, kappa)
        a = orc.to_dense(A.to_host(), N, b)
    finally:
        A.free()
        gen.free()
    assert np.array_equal(a, a.T)
    s = kappa ** (-np.arange(N) / (N - 1))
    ev = np.sort(np.linalg.eigvalsh(a))[::-1]
    assert np.abs(ev - s).max() < 50 * N * EPS
