"""CPU tests of the per-task trace (palial_amd.trace): the task names against
the reference's own recorded trace (tests/golden/cholesky_trace_names_mt16.txt,
from exp/data/tasks_times_8192_512_31_ondemand), the library's name list
against this module's, the DAG edges against hand-derived cases and the
dependence rules of src/cholesky.h / src/lu.h / src/qr.h, and the line format
(palial_trace/log.h:43-50) on synthetic records."""
import os

import numpy as np
import pytest

from palial_amd import trace

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cholesky_trace_names_mt16.txt")


def test_names_match_reference_trace():
    ref = [l.strip() for l in open(GOLDEN) if l.strip()]
    got = [name for name, _ in trace.cholesky_tasks(16)]
    assert len(ref) == 816 and got == ref


@pytest.mark.parametrize("algo", ["cholesky", "lu", "qr"])
@pytest.mark.parametrize("MT", [1, 2, 3, 7, 16])
def test_library_names_equal_module_names(algo, MT):
    import palial_amd
    assert palial_amd.task_names(algo, MT) == [n for n, _ in trace.tasks(algo, MT)]


def test_dag_edges_small_cases():
    def named(algo, MT):
        tl = trace.tasks(algo, MT)
        return {tl[i][0]: [tl[j][0] for j in d] for i, d in trace.dag_edges(algo, MT).items()}
    assert named("cholesky", 2) == {"potrf000": [], "trsm010": ["potrf000"], "syrk010": ["trsm010"],
                                    "potrf121": ["syrk010"]}
    assert named("lu", 2) == {"getrf000": [], "trsm010": ["getrf000"], "trsm021": ["getrf000"],
                              "gemm011": ["trsm010", "trsm021"], "getrf122": ["gemm011"]}
    # tsqrt(0,1) rewrites R in A(0,0) after ormqr read V there (depend inout vs in, src/qr.h:75,118)
    assert named("qr", 2) == {"geqrt000": [], "ormqr001": ["geqrt000"], "tsqrt012": ["geqrt000", "ormqr001"],
                              "tsmqr011": ["ormqr001", "tsqrt012"], "geqrt122": ["tsmqr011"]}


@pytest.mark.parametrize("MT", [3, 6])
def test_cholesky_gemm_dependences(MT):
    """gemm(k,m,n) updates A(n,m) with A(k,n) and A(k,m): it follows trsm(k,n),
    trsm(k,m) and the previous step's update of A(n,m), and nothing else."""
    tl = trace.tasks("cholesky", MT)
    idx = {t: i for i, (_, t) in enumerate(tl)}
    edges = trace.dag_edges("cholesky", MT)
    for (_, t) in tl:
        if t[0] != "gemm":
            continue
        _, k, m, n = t
        want = {idx[("trsm", k, n)], idx[("trsm", k, m)]}
        if k > 0:
            want.add(idx[("gemm", k - 1, m, n)])
        assert set(edges[idx[t]]) == want, t
    # every task but the first potrf depends on something, and only on earlier tasks
    assert all(all(j < i for j in d) for i, d in edges.items())
    assert all(d for i, d in edges.items() if i > 0)


def test_trace_lines_from_records():
    N, b = 768, 256
    tl = trace.tasks("lu", 3)
    rec = []
    for i, (_, t) in enumerate(tl):
        code, k, m, n = trace._key(t)
        rec.append([7 if code is None else code, k, m, n, 0.25 * (i + 1), 0.0])
    lines = trace.trace_lines("lu", np.array(rec), N, b)
    assert lines[0] == f"{tl[0][0]}, 0.250000"
    assert len(lines) == len(tl)
    bad = np.array(rec)
    bad[3, 1] += 1
    with pytest.raises(ValueError):
        trace.trace_lines("lu", bad, N, b)
