"""GPU tests of the batched small factorizations (no reference counterpart;
the per-matrix semantics are the reference tile kernels': LAPACKE_dpotrf('U')
at src/cholesky.h:40 and the no-pivot dgetrf of src/srclu/dgetrf_no_pivot.h).
Oracles: numpy Cholesky (unique factor of an SPD matrix), a plain Doolittle
elimination for the no-pivot LU, numpy solves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def doolittle(a):
    """No-pivot LU in place (L unit lower strictly below, U upper)."""
    a = a.copy()
    n = a.shape[0]
    for k in range(n - 1):
        a[k + 1:, k] /= a[k, k]
        a[k + 1:, k + 1:] -= np.outer(a[k + 1:, k], a[k, k + 1:])
    return a


def check_batch(hip, n, kind, prec, seed=3):
    bt = hip.Batch(n, kind, precision=prec)
    try:
        bt.generate(seed)
        mats0, _ = bt.get()
        rng = np.random.default_rng(seed)
        rhs = [rng.standard_normal(m) for m in n]
        bt.set(rhs=rhs)
        rep = bt.factor(solve=True)
        mats, xs = bt.get()
    finally:
        bt.free()
    assert (bt.info == 0).all()
    tol = 1e-12 if prec == hip.PREC_FP64 else 2e-5
    for a0, f, x, b, k in zip(mats0, mats, xs, rhs, kind):
        a0 = a0.astype(np.float64)
        m = a0.shape[0]
        if k == 0:
            u = np.linalg.cholesky(a0).T
            got = np.triu(f.astype(np.float64))
            assert np.abs(got - u).max() <= tol * np.abs(u).max()
            # the strict lower triangle keeps the input
            assert np.array_equal(np.tril(f, -1), np.tril(a0.astype(f.dtype), -1))
        else:
            lu = doolittle(a0)
            assert np.abs(f.astype(np.float64) - lu).max() <= tol * np.abs(lu).max()
        xe = np.linalg.solve(a0, b.astype(np.float64))
        assert np.abs(x.astype(np.float64) - xe).max() <= 10 * tol * max(1.0, np.abs(xe).max()), (m, k)
    return rep


@pytest.mark.parametrize("prec", [0, 1])
def test_batched_ragged(hip, prec):
    rng = np.random.default_rng(7)
    n = np.concatenate([[1, 2, 31, 32, 33, 64, 96, 511, 512], rng.integers(16, 300, 40)]).astype(np.int32)
    kind = (np.arange(len(n)) % 2).astype(np.int32)
    check_batch(hip, n, kind, prec)


def test_batched_single_kind(hip):
    n = np.array([40, 200, 128, 17], dtype=np.int32)
    check_batch(hip, n, np.zeros(4, dtype=np.int32), 0)
    check_batch(hip, n, np.ones(4, dtype=np.int32), 0)


def test_batched_info_on_failure(hip):
    n = np.array([48, 48], dtype=np.int32)
    kind = np.zeros(2, dtype=np.int32)
    bt = hip.Batch(n, kind)
    try:
        a = np.eye(48) * 4.0
        bad = a.copy()
        bad[20, 20] = -1.0
        bt.set(mats=[a, bad], rhs=[np.ones(48), np.ones(48)])
        bt.factor(solve=False)
    finally:
        bt.free()
    assert bt.info[0] == 0 and bt.info[1] == 21


def test_batched_generator_deterministic(hip):
    n = np.array([33, 70], dtype=np.int32)
    kind = np.array([0, 1], dtype=np.int32)
    out = []
    for _ in range(2):
        bt = hip.Batch(n, kind)
        try:
            bt.generate(11)
            out.append(bt.get()[0])
        finally:
            bt.free()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    s = out[0][0]
    assert np.array_equal(s, s.T)                      # kind 0 is symmetric
    assert np.all(np.diag(out[0][1]) > 60)             # G + n I


def test_batched_graph_replay_bitwise(hip):
    """The persistent batched launch (one kernel, largest matrices first from
    a work counter) gives the same bits eagerly, from its captured graph, and
    on replays of that graph with new data in the same buffers."""
    rng = np.random.default_rng(13)
    n = rng.integers(8, 200, 64).astype(np.int32)
    kind = (np.arange(64) % 2).astype(np.int32)
    bt = hip.Batch(n, kind)
    outs = []
    try:
        for seed, no_graph in ((5, True), (5, False), (6, False), (5, False)):
            bt.generate(seed)
            bt.set(rhs=[np.ones(m) for m in n])
            rep = bt.factor(solve=True, no_graph=no_graph)
            assert rep.nlaunches == 1 and (bt.info == 0).all()
            outs.append(bt.get())
    finally:
        bt.free()
    for (m0, x0), (m1, x1) in ((outs[0], outs[1]), (outs[0], outs[3])):
        assert all(np.array_equal(a, b) for a, b in zip(m0, m1))
        assert all(np.array_equal(a, b) for a, b in zip(x0, x1))
    assert not all(np.array_equal(a, b) for a, b in zip(outs[0][0], outs[2][0]))
