"""The CU-policy energy harness on the GPU (no reference counterpart; the
GPU analogue of palial_power/server.c:69-84): each policy's process starts
the library with its CU masks and solves correctly, and a malformed policy
is refused when the library starts, not silently ignored."""
import json

import pytest

pytestmark = pytest.mark.gpu


def run(capsys, argv):
    from palial_amd import energy
    rc = energy.main(argv)
    lines = [json.loads(x) for x in capsys.readouterr().out.splitlines() if x.startswith("{")]
    return rc, lines


def test_policies_solve(capsys):
    rc, lines = run(capsys, ["--n", "2048", "--tile", "256", "--runs", "2",
                             "--policies", "default,all_cus,chain_reserved,trsm_low,update_low"])
    assert rc == 0, lines
    assert [x["policy"] for x in lines] == ["default", "all_cus", "chain_reserved", "trsm_low", "update_low"]
    for x in lines:
        assert x["ms_median"] > 0 and x["runs"] == 2
        assert x["sensor"] is not None or x["energy_j"] is None


def test_bad_policy_refused(capsys):
    rc, lines = run(capsys, ["--n", "1024", "--tile", "256", "--runs", "1", "--policies", "bad=bulk=9-3"])
    assert rc == 1 and "error" in lines[0]
