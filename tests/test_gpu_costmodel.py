"""GPU tests of the cost model's HIP kernels (forward, gradient, training)
against the FP64 reference pinned in tests/test_costmodel.py, of the
launch enumeration against the launches the library records, and of one
small autotuning run."""
import numpy as np
import pytest

from palial_amd import costmodel as cm

pytestmark = pytest.mark.gpu


def test_forward_matches_reference(hip):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((300, cm.NFEAT)).astype(np.float32)
    m = cm.MLP(seed=2)
    got = m.predict_log_ms(X).astype(np.float64)
    ref = cm.forward_ref(m.params, X.astype(np.float64))
    assert np.abs(got - ref).max() <= 1e-4 * max(1.0, np.abs(ref).max())


def test_gradient_matches_reference(hip):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((333, cm.NFEAT)).astype(np.float32)     # two full chunks + a ragged one
    y = rng.standard_normal(333).astype(np.float32)
    m = cm.MLP(seed=4)
    g = m.gradient(X, y).astype(np.float64)
    ref = cm.grad_ref(m.params, X.astype(np.float64), y.astype(np.float64))
    assert np.abs(g - ref).max() <= 1e-4 * np.abs(ref).max()
    assert np.array_equal(g, m.gradient(X, y))                        # deterministic


def test_training_fits_a_smooth_target(hip):
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (1000, cm.NFEAT)).astype(np.float32)
    y = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3]).astype(np.float32)
    m = cm.MLP(seed=5)
    loss = m.train(X, y, epochs=150, lr=3e-3)
    assert loss[-1] < 0.1 * loss[0]
    pred = m.predict_log_ms(X)
    assert np.mean((pred - y) ** 2) < 0.2 * np.var(y)


@pytest.mark.parametrize("algo,prec", [("cholesky", 0), ("cholesky", 2), ("lu", 2)])
def test_enumeration_matches_recorded_launches(hip, orc, algo, prec):
    N, b = 1024, 256
    a = orc.gen("sym" if algo == "cholesky" else "dense", N, b, seed=1, shift=N)
    B = np.random.default_rng(0).standard_normal((N, 1))
    hip.profile(2)
    _, rep, rc = hip.solve(a, B, algo=algo, precision=prec, b=b, factor_tile=b)
    rec = hip.profile_records()
    hip.profile(0)
    assert rc == 0
    got = sorted((int(c), round(f), round(by)) for c, bb, p, f, by, ms in rec)
    exp = sorted((c, round(f), round(by)) for (_, _, c, f, by) in cm.launches(algo, N, b, prec))
    assert got == exp
    assert np.all(rec[:, 1] == b) and np.all(rec[:, 2] == prec) and np.all(rec[:, 5] >= 0)


def test_autotune_small(hip):
    from palial_amd import autotune
    N = 2048
    A = hip.DeviceMatrix(N, 256)
    try:
        A1, A2 = hip.kron_spd_factors(N, 1e2, seed=0)
        A.generate_kron(A1, A2)
        res = autotune.tune(hip, A, N, "cholesky", tiles=(256, 512), tiers=(0, 2), epochs=100)
    finally:
        A.free()
    assert res["best"] is not None and res["best"]["converged"]
    assert len(res["candidates"]) == 4
    assert res["train_loss"][1] < res["train_loss"][0]
