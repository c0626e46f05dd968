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


def test_features():
    keys = cm.task_keys("cholesky", 4)
    X = cm.features("cholesky", 2048, 512, 2, keys)
    assert X.shape == (len(keys), cm.NFEAT) and X.dtype == np.float32 and cm.NFEAT == 16
    assert (X[:, :cm.NTYPE].sum(1) == 1).all()
    assert (X[:, cm.NTYPE + 1:cm.NTYPE + 6].sum(1) == 1).all() and (X[:, cm.NTYPE + 3] == 1).all()
    # potrf000 is a diagonal task of depth 0 in a one-task launch, on the critical path
    assert X[0, 0] == 1 and X[0, cm.NTYPE + 7] == 0 and X[0, cm.NTYPE + 8] == 1 / 12 and X[0, cm.NTYPE + 9] == 1
    # chain: potrf(k), trsm(k,k+1), syrk(k,k+1) -- 3 per step but the last step's potrf only
    assert X[:, cm.NTYPE + 9].sum() == 3 * 3 + 1


def test_launch_tasks_cover_the_default_schedule():
    """The launch sizes of the split chain schedule: per step k the chain's
    one-tile update of (k+1,k+1), the first update launch (tile row k+1 and
    tile (k+2,k+2); LU also column k+1) and the rest: exactly `size` tasks
    carry each size, and the three add up to the step's trailing tiles."""
    MT = 7
    for algo, a in (("cholesky", 0), ("lu", 1)):
        keys = cm.task_keys(algo, MT)
        lt = cm.launch_tasks(a, MT, *keys.T)
        kind, k, m, n = keys.T
        up = (kind == 2) | (kind == 3) | (kind == 7)
        for kk in range(MT - 3):          # T >= 3: three distinct launch sizes
            T = MT - kk - 1
            sizes = lt[up & (k == kk)]
            first = T if a == 0 else 2 * T - 1
            rest = T * (T + 1) // 2 - 1 - T if a == 0 else (T - 1) ** 2 - 1
            assert sorted(set(sizes.tolist())) == sorted({1, first, rest})
            assert (sizes == 1).sum() == 1 and (sizes == first).sum() == first and (sizes == rest).sum() == rest
        trs = (kind == 1) | (kind == 5) | (kind == 6)
        assert set(lt[trs & (k == 0)].tolist()) == {1, MT - 2}


def test_replay_closed_forms(monkeypatch):
    monkeypatch.setattr(cm, "LAUNCH_MS", 0.0)     # dispatch gaps off: pure task times
    N, b = 4096, 512
    MT = N // b
    for algo in ("cholesky", "lu"):
        keys = cm.task_keys(algo, MT)
        diag = (keys[:, 0] == 0) | (keys[:, 0] == 4)
        ms = np.where(diag, 2.0, 0.0)
        for sched in (1, 2, 3, 4):      # only the diagonal tasks cost: MT of them in a row
            assert cm.replay(algo, N, b, 0, sched, ms, keys) == pytest.approx(2.0 * MT)
        rng = np.random.default_rng(1)
        ms = rng.uniform(0.1, 1.0, len(keys))
        one = cm.replay(algo, N, b, 0, 4, ms, keys)
        for sched in (1, 2):
            assert cm.replay(algo, N, b, 0, sched, ms, keys) <= one + 1e-9
    keys = cm.task_keys("qr", MT)
    ms = np.random.default_rng(2).uniform(0.1, 1.0, len(keys))
    t = [cm.replay("qr", N, b, 0, s, ms, keys) for s in (1, 2, 4, 8)]
    assert all(x >= y - 1e-9 for x, y in zip(t, t[1:]))
    # one chain: every panel task in a row, so at least their sum
    panel = (keys[:, 0] == 8) | (keys[:, 0] == 10)
    assert t[0] >= ms[panel].sum() - 1e-9


def test_calibrate():
    assert cm.calibrate([1.0, 2.0, 4.0], [3.0, 6.0, 12.0]) == pytest.approx(3.0)


def test_subset_design():
    from palial_amd import autotune as at
    c = at.factor_candidates("cholesky", 8192)
    assert {b for b, _ in c} == {128, 256, 512, 1024}
    sub = at.subset(c, 512, 1)
    assert len(sub) == 4 + 3 and all(b == 512 or s == 1 for b, s in sub)
    assert {b for b, _ in at.factor_candidates("cholesky", 24576)} == {128, 256, 384, 512, 768, 1024}
    assert {b for b, _ in at.factor_candidates("qr", 8192)} == {128, 256, 512}
    sc = at.solve_candidates(24576)
    assert (384, 3) in sc and (384, 4) in sc and all(t < 3 or b % 128 == 0 for b, t in sc)


def test_tune_knobs_candidates_and_restore():
    """The solve tuner's last stage: every admissible sweep block, paired
    updates tried for BF16 only, non-converged or fallen-back runs ignored,
    and the library's pairing setting restored."""
    from palial_amd import autotune

    assert autotune.sweep_blocks(8192, 512) == [1, 2, 4]
    assert autotune.sweep_blocks(1536, 512) == [1]          # 3 tiles: no block fits twice
    assert autotune.sweep_blocks(32768, 512, nrhs=2) == [1, 2]

    class FakeP:
        def __init__(self):
            self.pair = 1
            self.calls = []

        def set_option(self, name, v):
            assert name == "pair"
            prev, self.pair = self.pair, v
            return prev

    class FakeBench:
        def __init__(self, P):
            self.P = P

        def run(self, b, tier, reps=3, sweep_block=0):
            self.P.calls.append((sweep_block, self.P.pair))
            ms = 100.0 - 10 * sweep_block - 5 * self.P.pair
            return ms, None, {"converged": not (sweep_block == 4 and self.P.pair == 1), "tier_used": tier}

    P = FakeP()
    best, res = autotune.tune_knobs(P, FakeBench(P), 512, 2, 8192)
    assert sorted(P.calls) == sorted((q, pr) for q in (1, 2, 4) for pr in (0, 1))
    assert (4, 1) not in res and best == (4, 0)
    assert P.pair == 1
    P = FakeP()
    best, res = autotune.tune_knobs(P, FakeBench(P), 512, 0, 8192)
    assert {pr for _, pr in P.calls} == {0} and best == (4, 0)
    P = FakeP()
    autotune.tune_knobs(P, FakeBench(P), 512, 4, 8192)
    assert {pr for _, pr in P.calls} == {0, 1}
