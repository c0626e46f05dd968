"""Paired trailing updates of the BF16 and split-BF16 Cholesky and LU solves (no reference
counterpart): steps j and j+1 applied to the tiles n >= j+3 in one pass must
give a factor as good as the step-by-step updates, seen through the
refinement: same tier, about the same number of refinement steps, backward
error 1e-14 recomputed on the host, and the same solution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def backward_error(A, X, B):
    R = B - A @ X
    return max(np.abs(R[:, j]).max() / (np.abs(A).sum(1).max() * np.abs(X[:, j]).max() + np.abs(B[:, j]).max())
               for j in range(B.shape[1]))


@pytest.mark.parametrize("prec", [2, 4])
@pytest.mark.parametrize("algo", ["cholesky", "lu"])
@pytest.mark.parametrize("N,b,kappa", [(1280, 128, 0), (2048, 256, 0), (1536, 128, 1e3)])
def test_paired_updates_match(hip, orc, algo, N, b, kappa, prec):
    if kappa:
        r = np.random.default_rng(21)
        Q, _ = np.linalg.qr(r.standard_normal((N, N)))
        A = (Q * np.geomspace(1.0, 1.0 / kappa, N)) @ Q.T
        A = (A + A.T) / 2
        a = orc.from_dense(A, N, b)
    else:
        a = orc.gen("sym" if algo == "cholesky" else "dense", N, b, seed=22, shift=N)
        A = orc.to_dense(a, N, b)
    B = np.random.default_rng(23).standard_normal((N, 1))
    out = {}
    prev = hip.set_option("pair", 0)
    try:
        for pair in (0, 1):
            hip.set_option("pair", pair)
            X, rep, rc = hip.solve(a, B, algo=algo, precision=prec, b=b, factor_tile=b)
            assert rc == 0, (pair, rc, rep.as_dict())
            assert rep.precision_used == prec and rep.shifts == 0, (pair, rep.as_dict())
            assert backward_error(A, X, B) <= 2e-14, pair
            out[pair] = (X, rep.iterations)
    finally:
        hip.set_option("pair", prev)
    assert abs(out[0][1] - out[1][1]) <= 1, (out[0][1], out[1][1])
    X0, X1 = out[0][0], out[1][0]
    assert np.abs(X1 - X0).max() / np.abs(X0).max() < (1e-9 if kappa else 1e-11)


def test_set_option_rejects(hip):
    with pytest.raises(hip.PalialError):
        hip.set_option("pair", 2)
    with pytest.raises(hip.PalialError):
        hip.set_option("no_such_option", 1)
    prev = hip.set_option("sweep_block", 4)
    assert hip.set_option("sweep_block", prev) == 4
