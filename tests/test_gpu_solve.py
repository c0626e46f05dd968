"""GPU tests of the solve path (no reference counterpart: the reference only
factors).  The FP64 oracle here is numpy's dense solve and the closed-form
backward error; every precision tier must reach backward error <= 1e-14."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spd_cond(N, kappa, seed):
    """A = Q diag(s) Q^T, Q from the QR of a seeded Gaussian, s geometric from 1 to 1/kappa."""
    r = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(r.standard_normal((N, N)))
    s = np.geomspace(1.0, 1.0 / kappa, N)
    return (Q * s) @ Q.T


def backward_error(A, X, B):
    R = B - A @ X
    return max(np.abs(R[:, j]).max() / (np.abs(A).sum(1).max() * np.abs(X[:, j]).max() + np.abs(B[:, j]).max())
               for j in range(B.shape[1]))


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
@pytest.mark.parametrize("prec", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("N,b,nrhs", [(512, 128, 1), (1024, 256, 16), (768, 128, 3)])
def test_solve_diag_dominant(hip, orc, algo, prec, N, b, nrhs):
    a = orc.gen("sym" if algo == "cholesky" else "dense", N, b, seed=1, shift=N)
    A = orc.to_dense(a, N, b)
    B = np.random.default_rng(2).standard_normal((N, nrhs))
    X, rep, rc = hip.solve(a, B, algo=algo, precision=prec, b=b)
    assert rc == 0, (rc, rep.backward_error)
    assert rep.backward_error <= 1e-14
    assert backward_error(A, X, B) <= 2e-14
    assert np.abs(X - np.linalg.solve(A, B)).max() / np.abs(X).max() < 1e-11
    if prec == 0:
        assert rep.iterations <= 1
    if prec == 3:                              # the FP8 tier converges by itself here
        assert rep.precision_used == 3, rep.as_dict()


@pytest.mark.parametrize("prec,kappa", [(1, 1e2), (1, 1e5), (2, 1e2), (3, 1e2), (4, 1e2), (4, 1e5), (0, 1e8)])
def test_solve_conditioned_spd(hip, orc, prec, kappa):
    N, b = 512, 128
    A = spd_cond(N, kappa, 5)
    a = orc.from_dense(A, N, b)
    B = np.random.default_rng(6).standard_normal((N, 1))
    X, rep, rc = hip.solve(a, B, algo="cholesky", precision=prec, b=b, max_iter=60)
    assert rc == 0, (rc, rep.backward_error, rep.iterations)
    assert backward_error(A, X, B) <= 2e-14


def test_solve_leaves_A_untouched(hip, orc):
    N, b = 256, 64
    a = orc.gen("sym", N, b, seed=3, shift=N)
    keep = a.copy()
    hip.solve(a, np.ones((N, 1)), algo="cholesky", precision=1, b=b)
    assert np.array_equal(a, keep)


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
@pytest.mark.parametrize("b,ft", [(256, 128), (128, 256), (256, 64)])
def test_solve_factor_tile_relayout(hip, orc, algo, b, ft):
    """The solve factors a re-laid-out copy with its own tile size."""
    N = 1024
    a = orc.gen("sym" if algo == "cholesky" else "dense", N, b, seed=4, shift=N)
    A = orc.to_dense(a, N, b)
    B = np.random.default_rng(7).standard_normal((N, 2))
    for prec in (0, 2):
        X, rep, rc = hip.solve(a, B, algo=algo, precision=prec, b=b, factor_tile=ft)
        assert rc == 0 and backward_error(A, X, B) <= 2e-14


def gen_cond(N, kappa, seed, sym):
    """sym: A = Q diag(s) Q^T, Q from the QR of a seeded Gaussian.  Otherwise
    A = L0 diag(s) U0 with unit triangular L0, U0 (small random off-diagonal
    parts, well conditioned): its no-pivot LU is L0 and diag(s) U0, so the
    factorization without pivoting is stable and cond(A) ~ kappa.  s is
    geometric from 1 to 1/kappa, in a seeded random order."""
    r = np.random.default_rng(seed)
    s = r.permutation(np.geomspace(1.0, 1.0 / kappa, N))
    if sym:
        Q, _ = np.linalg.qr(r.standard_normal((N, N)))
        return (Q * s) @ Q.T
    L0 = np.eye(N) + np.tril(r.standard_normal((N, N)), -1) * (0.5 / np.sqrt(N))
    U0 = np.eye(N) + np.triu(r.standard_normal((N, N)), 1) * (0.5 / np.sqrt(N))
    return (L0 * s) @ U0


@pytest.mark.parametrize("algo,prec,kappa", [("cholesky", 1, 1e6), ("cholesky", 2, 1e3), ("lu", 1, 1e6),
                                             ("lu", 2, 1e3)])
@pytest.mark.parametrize("refine", [1, 0])
def test_gmres_ir_ill_conditioned(hip, orc, algo, prec, kappa, refine):
    """GMRES-IR (refine=1) and the automatic switch (refine=0) reach 1e-14
    where the low-precision factor alone contracts slowly or not at all."""
    N, b = 512, 128
    A = gen_cond(N, kappa, 8, algo == "cholesky")
    a = orc.from_dense(A, N, b)
    B = np.random.default_rng(9).standard_normal((N, 1))
    X, rep, rc = hip.solve(a, B, algo=algo, precision=prec, b=b, max_iter=40, refine=refine)
    assert rc == 0, (rc, rep.backward_error, rep.iterations, rep.gmres_iterations)
    assert backward_error(A, X, B) <= 2e-14
    if refine == 1:
        assert rep.gmres_iterations >= rep.iterations
    assert np.abs(X - np.linalg.solve(A, B)).max() <= kappa * 1e-13 * np.abs(X).max()


def test_gmres_ir_well_conditioned(hip, orc):
    N, b = 512, 128
    a = orc.gen("dense", N, b, seed=4, shift=N)
    A = orc.to_dense(a, N, b)
    B = np.random.default_rng(2).standard_normal((N, 2))
    X, rep, rc = hip.solve(a, B, algo="lu", precision=2, b=b, refine=1)
    assert rc == 0 and rep.iterations <= 3 and rep.gmres_iterations >= 1
    assert backward_error(A, X, B) <= 2e-14
    # LU-IR only: never uses GMRES
    X2, rep2, rc2 = hip.solve(a, B, algo="lu", precision=2, b=b, refine=2)
    assert rc2 == 0 and rep2.gmres_iterations == 0


def test_split_bf16_needs_working_copy(hip, orc):
    """The split-BF16 tier runs only on the FP32 working copy: a factor tile that
    is not a multiple of 128 is refused with an argument error."""
    N, b = 256, 64
    a = orc.gen("sym", N, b, seed=1, shift=N)
    B = np.ones((N, 1))
    with pytest.raises(hip.PalialError):
        hip.solve(a, B, algo="cholesky", precision=hip.PREC_BF16X3, b=b)


def test_split_bf16_fewer_refinement_steps(hip, orc):
    """Split-BF16 products (~2^-16) leave a factor accurate enough that LU-IR
    needs fewer steps than with plain BF16 at the same condition number."""
    N, b = 1024, 256
    A = spd_cond(N, 1e3, 9)
    a = orc.from_dense(A, N, b)
    B = np.random.default_rng(3).standard_normal((N, 1))
    its = {}
    for prec in (2, 4):
        X, rep, rc = hip.solve(a, B, algo="cholesky", precision=prec, b=b, refine=2, max_iter=60)
        assert rc == 0 and backward_error(A, X, B) <= 2e-14
        its[prec] = rep.iterations
    assert its[4] < its[2], its


@pytest.mark.parametrize("prec,kappa", [(1, 1e9), (2, 1e7), (4, 1e9)])
def test_breakdown_shift_or_fallback(hip, orc, prec, kappa):
    """A low-precision Cholesky of an ill-conditioned SPD matrix meets a
    non-positive pivot.  With no_fallback the solve reports it (rc = the
    column, > 0); by default it refactors A + sigma |diag A| (GMRES-IR after
    that) or solves in FP64, and the answer reaches backward error 1e-14."""
    N, b = 1024, 256
    A = spd_cond(N, kappa, 11)
    a = orc.from_dense(A, N, b)
    B = np.random.default_rng(4).standard_normal((N, 1))
    X0, rep0, rc0 = hip.solve(a, B, algo="cholesky", precision=prec, b=b, check=False, no_fallback=True)
    X, rep, rc = hip.solve(a, B, algo="cholesky", precision=prec, b=b, check=False)
    info = (rc0, rc, {k: rep.as_dict()[k] for k in ("precision_used", "shifts", "iterations", "gmres_iterations",
                                                    "backward_error")})
    assert rc == 0 and backward_error(A, X, B) <= 1e-14, (info, backward_error(A, X, B))
    if rc0 > 0:                                   # the plain factorization broke down
        assert rep.shifts > 0 or rep.precision_used == 0, info
    else:
        assert rep.shifts == 0, info
    assert rep.precision_used in (0, prec), info


def test_lu_nonconvergence_falls_back(hip, orc):
    """BF16 LU-IR on kappa = 1e9 does not converge in 5 steps: the default
    solve finishes in FP64 (precision_used 0); no_fallback reports it."""
    N, b = 1024, 256
    A = spd_cond(N, 1e9, 12)
    a = orc.from_dense(A, N, b)
    B = np.random.default_rng(5).standard_normal((N, 1))
    X0, rep0, rc0 = hip.solve(a, B, algo="lu", precision=2, b=b, max_iter=5, refine=2, check=False, no_fallback=True)
    assert rc0 != 0
    X, rep, rc = hip.solve(a, B, algo="lu", precision=2, b=b, max_iter=5, refine=2)
    assert rc == 0 and rep.precision_used == 0 and backward_error(A, X, B) <= 1e-14
