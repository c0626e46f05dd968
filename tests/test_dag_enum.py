"""CPU tests of the batched-launch task enumeration (host view of the device
decode, via the palial_hip_task_tiles test hook): every tile index a launch
touches is inside the N x N buffer, and the launches of one DAG step cover
exactly the reference's tile tasks (src/cholesky.h:107-173, src/lu.h:146-172,
src/qr.h:61-185) with the reference's operands."""
import ctypes as C

import pytest

CHOL_TRAIL, LU_TRAIL, QR_W, QR_TW, QR_VW, ORM_W, ORM_TW, ORM_VW = 1, 2, 3, 4, 5, 6, 7, 8
CHOL_ROW, CHOL_REST, LU_PANEL, LU_REST = 9, 10, 11, 12


def tasks(mode, MT, k, m=0, b=4):
    import palial_amd
    L = palial_amd.lib()
    L.palial_hip_task_tiles.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_long), C.c_int]
    cap = 3 * MT * MT + 3
    buf = (C.c_long * cap)()
    n = L.palial_hip_task_tiles(mode, MT, b, k, m, buf, cap)
    assert n >= 0
    out = []
    for t in range(n):
        trip = []
        for off in buf[3 * t:3 * t + 3]:
            if off < 0:
                trip.append(off)
                continue
            assert off % (b * b) == 0 and 0 <= off <= (MT * MT - 1) * b * b, (mode, MT, k, off)
            idx = off // (b * b)
            trip.append((idx % MT, idx // MT))      # (row tile, column tile)
        out.append(tuple(trip))
    return out


@pytest.mark.parametrize("MT", [1, 2, 3, 5, 8])
def test_cholesky_step_tasks(MT):
    for k in range(MT - 1):
        full = tasks(CHOL_TRAIL, MT, k)
        expect = {((k, n), (k, m), (n, m)) for m in range(k + 1, MT) for n in range(k + 1, m + 1)}
        assert len(full) == len(expect) and set(full) == expect
        row, rest = tasks(CHOL_ROW, MT, k), (tasks(CHOL_REST, MT, k) if MT - k - 1 > 1 else [])
        assert set(row) | set(rest) == expect and not set(row) & set(rest)
        assert all(t[2][0] == k + 1 for t in row) and all(t[2][0] >= k + 2 for t in rest)


@pytest.mark.parametrize("MT", [1, 2, 3, 5, 8])
def test_lu_step_tasks(MT):
    for k in range(MT - 1):
        full = tasks(LU_TRAIL, MT, k)
        expect = {((m, k), (k, n), (m, n)) for m in range(k + 1, MT) for n in range(k + 1, MT)}
        assert len(full) == len(expect) and set(full) == expect
        pan = tasks(LU_PANEL, MT, k)
        rest = tasks(LU_REST, MT, k) if MT - k - 1 > 1 else []
        assert set(pan) | set(rest) == expect and not set(pan) & set(rest)
        assert all(t[2][0] >= k + 2 and t[2][1] >= k + 2 for t in rest)


@pytest.mark.parametrize("MT", [2, 4, 6])
def test_qr_step_tasks(MT):
    for k in range(MT - 1):
        T = MT - k - 1
        for m in range(k + 1, MT):
            w = tasks(QR_W, MT, k, m)
            assert [t[0] for t in w] == [(m, k)] * T                       # V = A(m,k)
            assert [t[1] for t in w] == [(m, n) for n in range(k + 1, MT)]  # A2 = A(m,n)
            assert all(t[2] == -2 for t in w)                              # W workspace
            vw = tasks(QR_VW, MT, k, m)
            assert [t[2] for t in vw] == [(m, n) for n in range(k + 1, MT)]
            tw = tasks(QR_TW, MT, k, m)
            assert all(t[0] == -2 for t in tw)                             # T = S(m,k)
        o = tasks(ORM_VW, MT, k)
        assert [t[2] for t in o] == [(k, n) for n in range(k + 1, MT)]
        assert all(t[0] == (k, k) for t in o)
