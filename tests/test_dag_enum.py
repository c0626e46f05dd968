"""CPU tests of the batched-launch task enumeration (host view of the device
decode, via the palial_hip_task_tiles test hook): every tile index a launch
touches is inside the N x N buffer, and the launches of one DAG step cover
exactly the reference's tile tasks (src/cholesky.h:107-173, src/lu.h:146-172,
src/qr.h:61-185) with the reference's operands."""
import ctypes as C

import pytest

CHOL_TRAIL, LU_TRAIL, QR_W, QR_TW, QR_VW, ORM_W, ORM_TW, ORM_VW = 1, 2, 3, 4, 5, 6, 7, 8
CHOL_ROW, CHOL_REST, LU_PANEL, LU_REST = 9, 10, 11, 12
CHOL_FIRST, LU_FIRST = 13, 14
CHOL_ROW2, CHOL_REST2, LU_ROW2, LU_REST2 = 23, 24, 25, 26


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


@pytest.mark.parametrize("MT", [3, 4, 7])
def test_split_update_tasks(MT):
    """Split trailing update of the chain schedule: the chain updates tile
    (k+1,k+1); launch FIRST (from task 1) the rest of tile row k+1 (LU: row and
    column k+1) plus tile (k+2,k+2); launch REST (from task 1) the remainder.
    Together they are step k's update, each tile once."""
    for k in range(MT - 2):
        chain = {(k + 1, k + 1)}
        for full_mode, first_mode, rest_mode in ((CHOL_TRAIL, CHOL_FIRST, CHOL_REST), (LU_TRAIL, LU_FIRST, LU_REST)):
            full = tasks(full_mode, MT, k)
            first = tasks(first_mode, MT, k)
            rest = tasks(rest_mode, MT, k)
            assert first[0][2] == (k + 1, k + 1) and rest[0][2] == (k + 2, k + 2)   # task 0: skipped by task0 = 1
            first, rest = first[1:], rest[1:]
            assert (k + 2, k + 2) in [t[2] for t in first]
            outs = [t[2] for t in first] + [t[2] for t in rest]
            assert len(outs) == len(set(outs)) and not chain & set(outs)
            assert set(first) | set(rest) | {t for t in full if t[2] in chain} == set(full)
            if full_mode == CHOL_TRAIL:
                assert all(t[2][0] == k + 1 or t[2] == (k + 2, k + 2) for t in first)


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


@pytest.mark.parametrize("MT", [5, 6, 9])
def test_paired_update_tasks(MT):
    """Paired steps j, j+1 (T = MT-j-1 >= 4): step j's update is the chain's
    tile (j+1,j+1), FIRST (from task 1), ROW2 (row j+2 and tile (j+3,j+3)) and
    the tiles of REST2 (from task 1); REST2's tiles are also exactly step j+1's
    REST (from task 1), so one REST2 launch applies both steps there."""
    for j in range(MT - 4):
        full = tasks(CHOL_TRAIL, MT, j)
        first = tasks(CHOL_FIRST, MT, j)[1:]
        row2 = tasks(CHOL_ROW2, MT, j)
        rest2 = tasks(CHOL_REST2, MT, j)
        assert rest2[0][2] == (j + 3, j + 3)
        rest2 = rest2[1:]
        outs = [t[2] for t in first] + [t[2] for t in row2] + [t[2] for t in rest2]
        assert len(outs) == len(set(outs)) and (j + 1, j + 1) not in outs
        assert set(first) | set(row2) | set(rest2) | {t for t in full if t[2] == (j + 1, j + 1)} == set(full)
        assert all(t[2][0] == j + 2 or t[2] == (j + 3, j + 3) for t in row2)
        nxt = tasks(CHOL_REST, MT, j + 1)[1:]
        assert [t[2] for t in nxt] == [t[2] for t in rest2]


@pytest.mark.parametrize("MT", [5, 6, 9])
def test_paired_lu_update_tasks(MT):
    """LU pairs: FIRST (from task 1), ROW2 (column and row j+2, tile (j+3,j+3))
    and REST2 (from task 1) with the chain's (j+1,j+1) are step j's update;
    REST2's tiles are step j+1's REST (from task 1)."""
    for j in range(MT - 4):
        full = tasks(LU_TRAIL, MT, j)
        first = tasks(LU_FIRST, MT, j)[1:]
        row2 = tasks(LU_ROW2, MT, j)
        rest2 = tasks(LU_REST2, MT, j)
        assert rest2[0][2] == (j + 3, j + 3)
        rest2 = rest2[1:]
        outs = [t[2] for t in first] + [t[2] for t in row2] + [t[2] for t in rest2]
        assert len(outs) == len(set(outs)) and (j + 1, j + 1) not in outs
        assert set(first) | set(row2) | set(rest2) | {t for t in full if t[2] == (j + 1, j + 1)} == set(full)
        assert all(j + 2 in t[2] or t[2] == (j + 3, j + 3) for t in row2)
        nxt = tasks(LU_REST, MT, j + 1)[1:]
        assert [t[2] for t in nxt] == [t[2] for t in rest2]
