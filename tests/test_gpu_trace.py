"""GPU tests of the per-task trace: the records of a profiled factorization
on the device map onto the reference's task list (names in creation order,
times adding up to the launches' time without packing)."""
import numpy as np
import pytest

from palial_amd import costmodel as cm
from palial_amd import trace

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 2])
def test_trace_of_a_profiled_solve(hip, orc, prec):
    N, b = 1024, 256
    a = orc.gen("sym", N, b, seed=1, shift=N)
    B = np.random.default_rng(0).standard_normal((N, 1))
    hip.profile(2)
    _, rep, rc = hip.solve(a, B, algo="cholesky", precision=prec, b=b, factor_tile=b)
    rec = hip.profile_records()
    hip.profile(0)
    assert rc == 0
    lines = trace.cholesky_trace_lines(rec, N, b, prec)
    names = [l.split(", ")[0] for l in lines]
    assert names == [n for n, _ in trace.cholesky_tasks(N // b)]
    vals = np.array([float(l.split(", ")[1]) for l in lines])
    packs = rec[rec[:, 0] == cm.CLASSES["pack"], 5].sum()
    assert (vals >= 0).all() and vals.sum() > 0
    assert abs(vals.sum() - (rec[:, 5].sum() - packs)) <= 1e-6 * len(vals) + 1e-9


def test_lu_trace_of_a_profiled_solve(hip, orc):
    N, b = 1024, 256
    a = orc.gen("dense", N, b, seed=1, shift=N)
    B = np.random.default_rng(0).standard_normal((N, 1))
    hip.profile(2)
    _, rep, rc = hip.solve(a, B, algo="lu", precision=2, b=b, factor_tile=b)
    rec = hip.profile_records()
    hip.profile(0)
    assert rc == 0
    lines = trace.trace_lines("lu", rec, N, b, 2)
    assert [l.split(", ")[0] for l in lines] == [n for n, _ in trace.lu_tasks(N // b)]
    vals = np.array([float(l.split(", ")[1]) for l in lines])
    packs = rec[rec[:, 0] == cm.CLASSES["pack"], 5].sum()
    assert (vals >= 0).all() and abs(vals.sum() - (rec[:, 5].sum() - packs)) <= 1e-6 * len(vals) + 1e-9


def test_qr_trace_of_a_profiled_factorization(hip):
    N, b = 1024, 256
    A, S = hip.DeviceMatrix(N, b), hip.DeviceMatrix(N, b)
    try:
        A.generate(hip.GEN_DENSE, seed=3)
        hip.profile(2)
        info, _ = hip.factor_device("qr", A, S=S)
        rec = hip.profile_records()
        hip.profile(0)
    finally:
        A.free()
        S.free()
    assert info == 0
    lines = trace.trace_lines("qr", rec, N, b)
    assert [l.split(", ")[0] for l in lines] == [n for n, _ in trace.qr_tasks(N // b)]
    vals = np.array([float(l.split(", ")[1]) for l in lines])
    assert (vals >= 0).all() and abs(vals.sum() - rec[:, 5].sum()) <= 1e-6 * len(vals) + 1e-9


def test_trace_cli_prints_reference_lines(capsys):
    assert trace.main(["-m", "1024", "-b", "256"]) == 0
    out = capsys.readouterr().out.strip().splitlines()
    assert [l.split(", ")[0] for l in out] == [n for n, _ in trace.cholesky_tasks(4)]
    assert all(float(l.split(", ")[1]) >= 0 for l in out)
    assert trace.main(["-a", "lu", "-m", "1024", "-b", "256"]) == 0
    out = capsys.readouterr().out.strip().splitlines()
    assert [l.split(", ")[0] for l in out] == [n for n, _ in trace.lu_tasks(4)]
    assert trace.main(["-a", "qr", "-m", "1024", "-b", "256"]) == 0
    out = capsys.readouterr().out.strip().splitlines()
    assert [l.split(", ")[0] for l in out] == [n for n, _ in trace.qr_tasks(4)]
