"""CPU test of the bench's variant runs: the child's JSON lines become one
entry per sweep block and pairing, its failure and its time limit become an
``error`` entry, and the child command carries the headline workload."""
import json
import subprocess
import types

import pytest

import bench


def args(**kw):
    a = dict(n=32768, b=512, kappa=1e2, seed=0)
    a.update(kw)
    return types.SimpleNamespace(**a)


def row(q, pair, ms):
    return json.dumps({"sweep_block": q, "pair": pair, "ms_median": ms, "factor_ms": ms - 10, "solve_ms": 5,
                       "iterations": 3, "backward_error": 1e-16, "rc": 0, "precision_used": 2})


def test_variants_parsed(monkeypatch):
    seen = {}

    def fake(cmd, **kw):
        seen["cmd"], seen["kw"] = cmd, kw
        out = "\n".join([row(1, 0, 70.5), "noise line", row(4, 1, 66.25)]) + "\n"
        return subprocess.CompletedProcess(cmd, 0, out, "")
    monkeypatch.setattr(subprocess, "run", fake)
    v = bench.run_variants(args(), "1:0,4:1")
    assert set(v) == {"q1_pair0", "q4_pair1"}
    assert v["q4_pair1"]["ms_median"] == 66.25 and v["q1_pair0"]["rc"] == 0
    cmd = seen["cmd"]
    assert cmd[1:3] == ["-m", "palial_amd.profile_solve"]
    opt = dict(zip(cmd[3::2], cmd[4::2]))
    assert opt == {"--algo": "cholesky", "--precision": "2", "--n": "32768", "--tile": "512", "--kappa": "100.0",
                   "--seed": "0", "--solves": "3", "--variants": "1:0,4:1"}
    assert seen["kw"]["timeout"] > 0


def test_variants_failure(monkeypatch):
    monkeypatch.setattr(subprocess, "run", lambda cmd, **kw: subprocess.CompletedProcess(
        cmd, 1, row(1, 0, 70.0) + "\n", "Traceback\nRuntimeError: solve returned 2\n"))
    v = bench.run_variants(args(), "1:0,2:0")
    assert v["error"] == "RuntimeError: solve returned 2" and "q1_pair0" in v

    def slow(cmd, **kw):
        raise subprocess.TimeoutExpired(cmd, kw["timeout"])
    monkeypatch.setattr(subprocess, "run", slow)
    assert "timed out" in bench.run_variants(args(), "1:0")["error"]

