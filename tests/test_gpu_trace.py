"""GPU tests of the per-task trace: every reference tile task of a
factorization gets a device-measured time (batched tasks from their own
workgroups' timestamps, diagonal and panel tasks from HIP events), the records
follow the reference's creation order and naming, and a tracer loaded in
front of the driver receives every task through the reference's hooks."""
import os
import subprocess

import numpy as np
import pytest

from palial_amd import trace

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("algo,N,b", [("cholesky", 1024, 256), ("lu", 1024, 256), ("qr", 1024, 256),
                                      ("cholesky", 2048, 128), ("qr", 768, 128)])
def test_every_task_measured(hip, algo, N, b):
    rec = trace.measure(algo, N, b, seed=1)
    lines = trace.trace_lines(algo, rec, N, b)           # names and kinds in creation order
    assert len(lines) == len(trace.tasks(algo, N // b))
    ms, start = rec[:, 4], rec[:, 5]
    assert (ms > 0).all(), f"{int((ms <= 0).sum())} tasks without a time"
    assert (ms < 1000).all() and (start > -1.0).all()
    # a task never starts before the tasks it depends on have started (starts
    # of batched tasks come from the device clock and those of diagonal and
    # panel tasks from HIP events: zero points a few microseconds apart).  The
    # schedule runs ormqr(k, n) beside tsqrt(k, m): they touch disjoint halves
    # of A(k,k) (V below, R above the diagonal), which the reference's tile
    # dependence does not distinguish.
    tl = trace.tasks(algo, N // b)
    edges = trace.dag_edges(algo, N // b)
    late = [(tl[i][0], tl[j][0]) for i, d in edges.items() for j in d
            if start[i] + 0.02 < start[j] and not (tl[i][1][0] == "tsqrt" and tl[j][1][0] == "ormqr")]
    assert not late, late[:5]


def test_mixed_solve_factorization_traced(hip):
    N, b = 1024, 256
    rec = trace.measure("cholesky", N, b, seed=2, precision=2)
    assert (rec[:, 4] > 0).all()
    trace.trace_lines("cholesky", rec, N, b)


def test_hooks_of_a_loaded_tracer(hip, tmp_path):
    src = os.path.join(os.path.dirname(__file__), "hook_tracer.c")
    so = tmp_path / "hook_tracer.so"
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", "-o", str(so), src], check=True)
    out = tmp_path / "hooks.txt"
    env = dict(os.environ, LD_PRELOAD=str(so), PALIAL_HOOK_OUT=str(out))
    for algo, N, b in (("cholesky", 1024, 256), ("qr", 512, 128)):
        r = subprocess.run([hip.DRIVER_PATH, "-a", algo, "-m", str(N), "-b", str(b), "-T"], env=env,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout == ""                                  # still silent on stdout
        rows = [l.split() for l in out.read_text().splitlines()]
        assert [x[0] for x in rows] == [n for n, _ in trace.tasks(algo, N // b)]
        assert all(int(x[1]) == 0 for x in rows)               # device 0
        assert all(int(x[3]) >= int(x[2]) for x in rows)
