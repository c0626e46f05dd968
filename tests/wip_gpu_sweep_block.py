"""Blocked correction sweeps (no reference counterpart: the reference only
factors).  Steps of q tiles apply precomputed inverses of the q x q tile
blocks on the diagonal; every q must give a solution of the same backward
error in about the same number of refinement steps as one-tile steps, and
blocks the sweep kernels cannot take are refused when asked for."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def backward_error(A, X, B):
    R = B - A @ X
    return max(np.abs(R[:, j]).max() / (np.abs(A).sum(1).max() * np.abs(X[:, j]).max() + np.abs(B[:, j]).max())
               for j in range(B.shape[1]))


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_sweep_blocks_agree(hip, orc, algo, prec, nrhs):
    N, b = 1024, 128
    a = orc.gen("sym" if algo == "cholesky" else "dense", N, b, seed=11, shift=N)
    A = orc.to_dense(a, N, b)
    B = np.random.default_rng(12).standard_normal((N, nrhs))
    its = {}
    for q in (1, 2, 4):
        X, rep, rc = hip.solve(a, B, algo=algo, precision=prec, b=b, sweep_block=q)
        assert rc == 0, (q, rc, rep.backward_error)
        assert rep.precision_used == prec
        assert backward_error(A, X, B) <= 2e-14, q
        assert np.abs(X - np.linalg.solve(A, B)).max() / np.abs(X).max() < 1e-11
        its[q] = rep.iterations
    assert max(its.values()) - min(its.values()) <= 1, its
    if prec == 0:
        assert max(its.values()) <= 1, its


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
def test_sweep_blocks_conditioned(hip, orc, algo):
    """Blocks of 4 tiles on a matrix with kappa ~ 1e5 (the inverses of the
    blocks are as ill-conditioned as the blocks): BF16 factors with GMRES-IR
    where needed still reach backward error 1e-14."""
    N, b = 1024, 128
    r = np.random.default_rng(13)
    Q, _ = np.linalg.qr(r.standard_normal((N, N)))
    s = np.geomspace(1.0, 1e-5, N)
    A = (Q * s) @ Q.T
    if algo == "lu":
        A = A + np.triu(r.standard_normal((N, N)) * 1e-3, 1)
    a = orc.from_dense(A, N, b)
    B = r.standard_normal((N, 1))
    for q in (1, 4):
        X, rep, rc = hip.solve(a, B, algo=algo, precision=2, b=b, sweep_block=q, max_iter=60)
        assert rc == 0, (q, rc, rep.as_dict())
        assert backward_error(A, X, B) <= 2e-14, q


@pytest.mark.parametrize("q,nrhs", [(3, 1), (16, 1), (2, 16), (8, 2)])
def test_sweep_block_refused(hip, orc, q, nrhs):
    """N / b = 8: q must divide 8 with 8 >= 2 q; at most 4 right-hand sides;
    q b (4 for 2..4 right-hand sides) <= 4096."""
    N, b = 1024, 128
    a = orc.gen("sym", N, b, seed=14, shift=N)
    B = np.ones((N, nrhs))
    _, _, rc = hip.solve(a, B, algo="cholesky", precision=0, b=b, sweep_block=q, check=False)
    assert rc == -1                            # PALIAL_ERR_ARG


def test_tuner_knobs(hip):
    """The solve tuner's last stage on the GPU: each admissible sweep block
    and pairing choice converges in the picked tier, and the best is taken."""
    from palial_amd import autotune
    N, b = 2048, 256
    gen = hip.CondGenerator(N, b, seed=0)
    A = hip.DeviceMatrix(N, b)
    sb = autotune.SolveBench(hip, A, "cholesky")
    try:
        gen.fill(A, 1e2)
        best, res = autotune.tune_knobs(hip, sb, b, 2, N, reps=2)
    finally:
        sb.free()
        A.free()
        gen.free()
    assert set(res) == {(q, p) for q in (1, 2, 4) for p in (0, 1)}, res
    assert best == min(res, key=res.get)
