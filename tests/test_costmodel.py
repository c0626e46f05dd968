"""CPU tests of the cost model's host side: the FP64 reference MLP (pinned by
finite differences, so the GPU kernels' checks against it mean something),
the featurization and the launch enumeration (flop totals against the closed
form of the tiled algorithms)."""
import numpy as np
import pytest

from palial_amd import costmodel as cm


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = cm.init_params(seed=3).astype(np.float64)
    X = rng.standard_normal((40, cm.NFEAT))
    y = rng.standard_normal(40)
    g = cm.grad_ref(p, X, y)
    eps = 1e-6
    for i in [0, 7, cm.NFEAT * 64 + 3, cm.NFEAT * 64 + 64 + 100, len(p) - 65, len(p) - 1]:
        q = p.copy(); q[i] += eps
        lp = np.mean((cm.forward_ref(q, X) - y) ** 2)
        q[i] -= 2 * eps
        lm = np.mean((cm.forward_ref(q, X) - y) ** 2)
        assert abs((lp - lm) / (2 * eps) - g[i]) <= 1e-6 * max(1.0, abs(g[i]))


def test_featurize():
    recs = np.array([[0, 512, 0, 2.7e11, 6e9, 3.4], [6, 256, 2, 1e9, 2e8, 0.1],
                     [6, 512, 3, 1e9, 2e8, 0.1], [6, 512, 4, 1e9, 2e8, 0.1]])
    X, y = cm.featurize(recs)
    assert X.shape == (4, cm.NFEAT) and X.dtype == np.float32
    assert X[0, 0] == 1 and X[1, 6] == 1 and X[0, cm.NCLASS + 3] == 1 and X[1, cm.NCLASS + 5] == 1
    # one tier column each: FP8 and split BF16 are not folded into BF16
    assert X[2, cm.NCLASS + 6] == 1 and X[3, cm.NCLASS + 7] == 1
    assert (X[:, cm.NCLASS + 3:].sum(1) == 1).all()
    assert np.allclose(y, np.log([3.4, 0.1, 0.1, 0.1]))


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
@pytest.mark.parametrize("prec", [0, 2])
def test_launch_flops_sum_to_tiled_count(algo, prec):
    N, b = 4096, 512
    MT = N // b
    L = cm.launches(algo, N, b, prec)
    tot = sum(f for (_, _, _, f, _) in L)
    b3 = float(b) ** 3
    pairs = MT * (MT - 1) / 2
    if algo == "cholesky":
        # POTRF + TRSM + SYRK + GEMM of the reference DAG (src/cholesky.h:32-173)
        exp = MT * b3 / 3 + pairs * b3 + pairs * b3 + MT * (MT - 1) * (MT - 2) / 6 * 2 * b3
    else:
        # GETRF + 2 TRSM per pair + GEMM over the trailing squares (src/lu.h)
        exp = MT * 2 * b3 / 3 + 2 * pairs * b3 + sum((MT - k - 1) ** 2 for k in range(MT)) * 2 * b3
    assert np.isclose(tot, exp, rtol=1e-12)
    assert {st for (_, st, _, _, _) in L} == {"chain", "aux", "first", "rest"}
    # one update launch per step: same flops
    tot1 = sum(f for (_, _, _, f, _) in cm.launches(algo, N, b, prec, split=False))
    assert np.isclose(tot1, exp, rtol=1e-12)
