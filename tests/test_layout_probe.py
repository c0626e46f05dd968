"""CPU test of the layout-L product helpers used by the large-N GPU checks,
against dense numpy on small matrices."""
import numpy as np
import pytest

from layout_probe import inf_norm, matvec


@pytest.mark.parametrize("N,b", [(48, 16), (64, 64), (96, 32)])
def test_matvec_parts(orc, N, b):
    a = orc.gen("dense", N, b, seed=2, shift=0.0)
    A = orc.to_dense(a, N, b)
    X = np.random.default_rng(1).standard_normal((N, 3))
    U = np.triu(A)
    L = np.tril(A, -1) + np.eye(N)
    for part, M in (("full", A), ("upper", U), ("upper_t", U.T), ("unit_lower", L)):
        assert np.allclose(matvec(a, N, b, X, part), M @ X, rtol=1e-13, atol=1e-13), part
    assert np.allclose(matvec(a, N, b, X[:, 0]), A @ X[:, 0])
    assert inf_norm(a, N, b) == pytest.approx(np.abs(A).sum(1).max(), rel=1e-14)
