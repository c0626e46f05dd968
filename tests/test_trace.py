"""CPU tests of the per-task trace (palial_amd.trace): the task names against
the reference's own recorded trace (tests/golden/cholesky_trace_names_mt16.txt,
from exp/data/tasks_times_8192_512_31_ondemand), the launch -> task mapping
against the task set of src/cholesky.h:14-173 (every task exactly once), and
the time split on synthetic per-launch records."""
import os

import numpy as np
import pytest

from palial_amd import costmodel as cm
from palial_amd import trace

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cholesky_trace_names_mt16.txt")


def test_names_match_reference_trace():
    ref = [l.strip() for l in open(GOLDEN) if l.strip()]
    got = [name for name, _ in trace.cholesky_tasks(16)]
    assert len(ref) == 816 and got == ref


@pytest.mark.parametrize("MT", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("split", [True, False])
def test_launches_cover_every_task_once(MT, split):
    b = 256
    tasks = [t for _, t in trace.cholesky_tasks(MT)]
    assert len(set(tasks)) == len(tasks)
    covered = []
    for (k, st, c, f, by) in cm.launches("cholesky", MT * b, b, 0, split=split):
        got = trace._covered(k, st, c, MT, split)
        covered += got
        if got:   # the launch's flops are its tasks' flops
            assert abs(sum(trace._flops(t, b) for t in got) - f) <= 1e-9 * f
    assert sorted(covered) == sorted(tasks)


@pytest.mark.parametrize("prec", [0, 2])
def test_times_split_over_tasks(prec):
    N, b = 2048, 256
    L = cm.launches("cholesky", N, b, prec)
    rng = np.random.default_rng(0)
    ms = rng.uniform(0.01, 1.0, len(L))
    rec = np.array([(c, b, prec, f, by, t) for (_, _, c, f, by), t in zip(L, ms)])
    times = trace.cholesky_task_times(rec, N, b, prec)
    packs = sum(t for (_, _, c, _, _), t in zip(L, ms) if c == cm.CLASSES["pack"])
    assert abs(sum(times.values()) - (ms.sum() - packs)) < 1e-9
    lines = trace.cholesky_trace_lines(rec, N, b, prec)
    assert len(lines) == len(trace.cholesky_tasks(N // b))
    name, val = lines[0].split(", ")
    assert name == "potrf000" and abs(float(val) - times[("potrf", 0)]) < 1e-6
    # the potrf of step 0 is a launch of its own
    diag0 = [t for (k, _, c, _, _), t in zip(L, ms) if k == 0 and c == cm.CLASSES["diag"]]
    assert abs(times[("potrf", 0)] - diag0[0]) < 1e-12


def test_partial_records_are_rejected():
    N, b = 1024, 256
    L = cm.launches("cholesky", N, b, 0)
    rec = np.array([(c, b, 0, f, by, 1.0) for (_, _, c, f, by) in L])
    with pytest.raises(ValueError):
        trace.cholesky_task_times(rec[:-1], N, b, 0)
    with pytest.raises(ValueError):
        trace.cholesky_task_times(np.vstack([rec, rec[:1]]), N, b, 0)


def test_lu_names_follow_the_reference_loops():
    """src/lu.h:13-173 by hand for 3 x 3 tiles: getrf, the column panel's trsm
    over m, then per n the row panel's trsm and the gemm column; m and n keep
    their last loop values (the reference leaves them uninitialised at k = 0)."""
    exp = ["getrf000", "trsm010", "trsm020", "trsm031", "gemm011", "gemm021", "trsm032", "gemm012", "gemm022",
           "getrf133", "trsm123", "trsm132", "gemm122", "getrf233"]
    assert [n for n, _ in trace.lu_tasks(3)] == exp


@pytest.mark.parametrize("MT", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("split", [True, False])
def test_lu_launches_cover_every_task_once(MT, split):
    b = 256
    tasks = [t for _, t in trace.lu_tasks(MT)]
    assert len(set(tasks)) == len(tasks)
    covered, occ = [], {}
    for (k, st, c, f, by) in cm.launches("lu", MT * b, b, 0, split=split):
        idx = occ.get((k, st, c), 0)
        occ[(k, st, c)] = idx + 1
        got = trace._covered_lu(k, st, c, MT, split, idx)
        covered += got
        if got:
            assert abs(sum(trace._flops(t, b) for t in got) - f) <= 1e-9 * f
    assert sorted(covered) == sorted(tasks)


def test_lu_times_add_up():
    N, b = 2048, 256
    L = cm.launches("lu", N, b, 2)
    ms = np.random.default_rng(1).uniform(0.01, 1.0, len(L))
    rec = np.array([(c, b, 2, f, by, t) for (_, _, c, f, by), t in zip(L, ms)])
    times = trace.task_times("lu", rec, N, b, 2)
    packs = sum(t for (_, _, c, _, _), t in zip(L, ms) if c == cm.CLASSES["pack"])
    assert abs(sum(times.values()) - (ms.sum() - packs)) < 1e-9
    assert len(trace.trace_lines("lu", rec, N, b, 2)) == len(trace.lu_tasks(N // b))


def test_qr_names_follow_the_reference_loops():
    """src/qr.h:17-185 by hand for 3 x 3 tiles: geqrt, ormqr over n, then per
    m tsqrt and the tsmqr row; m and n keep their last loop values."""
    exp = ["geqrt000", "ormqr001", "ormqr002", "tsqrt013", "tsmqr011", "tsmqr012", "tsqrt023", "tsmqr021",
           "tsmqr022", "geqrt133", "ormqr132", "tsqrt123", "tsmqr122", "geqrt233"]
    assert [n for n, _ in trace.qr_tasks(3)] == exp


@pytest.mark.parametrize("MT", [1, 2, 5])
def test_qr_times_cover_every_task(MT):
    N, b = MT * 256, 256
    L = trace.qr_launches(N, b)
    ms = np.random.default_rng(2).uniform(0.01, 1.0, len(L))
    rec = np.array([(c, b, 0, f, by, t) for (_, c, f, by), t in zip(L, ms)])
    times = trace.task_times("qr", rec, N, b)
    assert set(times) == {t for _, t in trace.qr_tasks(MT)}
    assert abs(sum(times.values()) - ms.sum()) < 1e-9
    with pytest.raises(ValueError):
        trace.task_times("qr", rec[:-1], N, b)
