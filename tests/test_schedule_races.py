"""CPU checks of the diagonal-chain schedules of the solves (no reference
counterpart: the reference's OpenMP tasks are ordered by their depend
clauses, src/cholesky.h:36-173 and src/lu.h:34-172; here four HIP streams
are ordered by events).  The library records one schedule on the host
(palial_hip_schedule_ops: every launch with the tiles, packed panel tiles and
diagonal-block inverses it reads and writes, every event record and wait)
and this test checks it as a race detector would:

- happens-before from stream order and event record -> wait (vector clocks);
  every two launches on different streams that touch one resource, one of
  them writing it, must be ordered;
- issue order, which every such order agrees with, must compute the
  factorization: each tile gets the update of every earlier step exactly
  once and before it is factored or solved, updates read panel tiles that
  were solved and packed for that step, and every tile ends solved.

FP64, FP32, BF16 (step by step and with paired updates), FP8 and split BF16
schedules of Cholesky and LU are covered, and the QR schedule with 1 to 8
panel chains in flight (src/qr.h:15-192), whose per-tile write order must
be the reference's and whose reads must see the reference's last writer."""
import ctypes as C

import numpy as np
import pytest

LAUNCH, RECORD, WAIT = 0, 1, 2
DIAG, TRSM, GEMM, PACK = 1, 2, 3, 4
RES = 1 << 24


def ops(algo, MT, b, prec, pair):
    import palial_amd
    L = palial_amd.lib()
    L.palial_hip_schedule_ops.argtypes = [C.c_int] * 5 + [C.POINTER(C.c_long), C.c_long]
    n = L.palial_hip_schedule_ops(algo, MT, b, prec, pair, None, 0)
    assert n > 0, n
    buf = (C.c_long * n)()
    assert L.palial_hip_schedule_ops(algo, MT, b, prec, pair, buf, n) == n
    v = np.array(buf[:], dtype=np.int64)
    out, i = [], 0
    while i < n:
        typ, st, ev, kind, mode, k, m, nr, nw, nx = (int(x) for x in v[i:i + 10])
        i += 10
        rd = set(v[i:i + nr].tolist())
        i += nr
        wr = set(v[i:i + nw].tolist())
        i += nw
        x = v[i:i + 4 * nx].reshape(nx, 4).tolist()
        i += 4 * nx
        out.append(dict(type=typ, stream=st, ev=ev, kind=kind, mode=mode, k=k, m=m, rd=rd, wr=wr, x=x))
    return out


def happens_before(seq):
    """Vector clocks: launch i gets (stream, position, clock); a -> b iff
    b.clock[a.stream] >= a.position."""
    clock, evc, pos = {}, {}, {}
    launches = []
    for o in seq:
        s = o["stream"]
        c = clock.setdefault(s, {})
        if o["type"] == RECORD:
            evc[o["ev"]] = dict(c)
        elif o["type"] == WAIT:
            for t, p in evc.get(o["ev"], {}).items():
                c[t] = max(c.get(t, -1), p)
        else:
            p = pos.get(s, 0)
            pos[s] = p + 1
            c[s] = p
            launches.append((o, s, p, dict(c)))
    return launches


def check_races(launches):
    n = 0
    for i, (a, sa, pa, ca) in enumerate(launches):
        for b_, sb, pb, cb in launches[i + 1:]:
            if sa == sb:
                continue
            if not (a["wr"] & (b_["rd"] | b_["wr"]) or b_["wr"] & a["rd"]):
                continue
            n += 1
            assert cb.get(sa, -1) >= pa or ca.get(sb, -1) >= pb, (a["kind"], sa, b_["kind"], sb,
                                                                 sorted((a["wr"] & (b_["rd"] | b_["wr"])) |
                                                                        (b_["wr"] & a["rd"]))[:4])
    return n


def check_semantics(seq, algo, MT):
    """Issue-order execution of the recorded operations, on tile states."""
    def rc(t):
        return t % MT, t // MT
    upd = {t: set() for t in range(MT * MT)}     # steps applied to each tile
    solved = set()
    packed = {}                                  # packed id -> (source tile, step)
    relevant = [(m, n) for m in range(MT) for n in range(MT) if algo == 1 or m <= n]
    for o in seq:
        if o["type"] != LAUNCH:
            continue
        if o["kind"] == DIAG:
            for t, k, _, _ in o["x"]:
                m, n = rc(t)
                assert m == n == k and upd[t] == set(range(k)) and t not in solved
                solved.add(t)
        elif o["kind"] == TRSM:
            for t, k, d, _ in o["x"]:
                m, n = rc(t)
                assert d == k + MT * k and d in solved and t not in solved, (m, n, k)
                assert min(m, n) == k and upd[t] == set(range(k)), (m, n, k, upd[t])
                solved.add(t)
        elif o["kind"] == PACK:
            for t, k, pid, _ in o["x"]:
                assert t in solved, rc(t)
                packed[pid] = (t, k)
        else:
            for c, k, a, bb in o["x"]:
                m, n = rc(c)
                assert c not in solved and k not in upd[c] and k < min(m, n), (m, n, k)
                # operands: Cholesky U(k, m), U(k, n); LU L(m, k), U(k, n)
                want = (k + MT * m, k + MT * n) if algo == 0 else (m + MT * k, k + MT * n)
                got = []
                for opnd in (a, bb):
                    if opnd >= RES:
                        assert opnd in packed, (m, n, k)
                        src, ks = packed[opnd]
                        assert ks == k
                        got.append(src)
                    else:
                        got.append(opnd)
                    assert got[-1] in solved
                assert tuple(got) == want, (m, n, k, [rc(g) for g in got])
                upd[c].add(k)
    for m, n in relevant:
        assert m + MT * n in solved, (m, n)


CASES = [(algo, MT, b, prec, variant) for algo in (0, 1) for (MT, b) in ((2, 128), (5, 128), (8, 128), (9, 256))
         for prec, variant in ((0, 0), (0, 2), (0, 4), (0, 6), (1, 0), (2, 0), (2, 1), (3, 0), (4, 0), (4, 1))]


@pytest.mark.parametrize("algo,MT,b,prec,variant", CASES)
def test_chain_schedule_race_free_and_complete(algo, MT, b, prec, variant):
    """variant = paired updates + 2 * schedule: 0 diagonal chain with the split
    update (every tier), 1 chain with one update launch, 2 row look-ahead,
    3 one stream (FP64)."""
    seq = ops(algo, MT, b, prec, variant)
    launches = happens_before(seq)
    assert launches and {o["stream"] for o, _, _, _ in launches} <= {1, 2, 3, 4}
    nconf = check_races(launches)
    assert nconf > 0 or MT < 3 or variant == 6    # two tiles, or one stream: nothing to order
    check_semantics(seq, algo, MT)


def test_pairing_changes_the_bulk_launches():
    """With pairing, steps j (even, >= 4 trailing tiles) and j+1 share one
    launch over the tiles n >= j+3: the BF16 schedule has fewer GEMM launches
    whose products come from two steps."""
    for algo in (0, 1):
        one = [o for o in ops(algo, 8, 128, 2, 0) if o["kind"] == GEMM]
        two = [o for o in ops(algo, 8, 128, 2, 1) if o["kind"] == GEMM]
        steps = [len({k for _, k, _, _ in o["x"]}) for o in two]
        assert max(steps) == 2 and max(len({k for _, k, _, _ in o["x"]}) for o in one) == 1
        assert sum(len(o["x"]) for o in one) == sum(len(o["x"]) for o in two)


def test_schedule_hook_rejects():
    import palial_amd
    L = palial_amd.lib()
    L.palial_hip_schedule_ops.argtypes = [C.c_int] * 5 + [C.POINTER(C.c_long), C.c_long]
    assert L.palial_hip_schedule_ops(2, 8, 128, 0, 0, None, 0) < 0
    assert L.palial_hip_schedule_ops(0, 8, 64, 2, 0, None, 0) < 0       # packed panels need b % 128 == 0
    assert L.palial_hip_schedule_ops(0, 8, 128, 2, 2, None, 0) < 0      # mixed tiers: the split chain only
    assert L.palial_hip_schedule_ops(2, 8, 128, 0, 9, None, 0) < 0      # at most 8 QR chains


QR_W, QR_TW, QR_VW, ORM_W, ORM_TW, ORM_VW = 3, 4, 5, 6, 7, 8
QRPANEL, VCOPY = 5, 6
WORKSPACE = 4 * RES     # W, op(T) W, tau and Gram workspace ids start here (S tiles below)


def qr_label(o):
    if o["kind"] == QRPANEL:
        return ("G", o["k"]) if o["mode"] == 0 else ("Q", o["k"], o["m"])
    if o["kind"] == VCOPY:        # ORMQR(k)'s copy of V(k), read by its products
        return ("O", o["k"])
    if o["mode"] in (ORM_W, ORM_TW, ORM_VW):
        return ("O", o["k"])
    assert o["mode"] in (QR_W, QR_TW, QR_VW)
    return ("T", o["k"], o["m"])


def qr_reference(MT):
    """The reference's loop order (src/qr.h:17-185) with the resources each
    task writes and reads (tiles m + MT n, V of diagonal tile k at 8 RES + ...,
    T-factor tiles at 3 RES + ...)."""
    def t(m, n):
        return m + MT * n
    def V(k):
        return 8 * RES + t(k, k)
    def S(m, n):
        return 3 * RES + t(m, n)
    ops = []
    for k in range(MT):
        ops.append((("G", k), {t(k, k), V(k), S(k, k)}, {t(k, k)}))
        if k + 1 == MT:
            break
        ops.append((("O", k), {t(k, n) for n in range(k + 1, MT)},
                    {V(k), S(k, k)} | {t(k, n) for n in range(k + 1, MT)}))
        for m in range(k + 1, MT):
            ops.append((("Q", k, m), {t(k, k), t(m, k), S(m, k)}, {t(k, k), t(m, k)}))
            ops.append((("T", k, m), {t(k, n) for n in range(k + 1, MT)} | {t(m, n) for n in range(k + 1, MT)},
                        {t(m, k), S(m, k)} | {t(k, n) for n in range(k + 1, MT)} | {t(m, n) for n in range(k + 1, MT)}))
    return ops


def writer_history(ops_):
    """Per resource: the labels of its writers in order; per (label, resource
    read): the label of the resource's last writer before it (None: input)."""
    hist, seen = {}, {}
    for lab, wr, rd in ops_:
        for r in rd:
            if r < WORKSPACE or r >= 8 * RES:
                seen.setdefault((lab, r), hist.get(r, [None])[-1])
        for r in wr:
            if r < WORKSPACE or r >= 8 * RES:
                h = hist.setdefault(r, [])
                if not h or h[-1] != lab:
                    h.append(lab)
    return hist, seen


@pytest.mark.parametrize("MT", [3, 5, 8])
@pytest.mark.parametrize("ns", [1, 2, 4, 8])
def test_qr_schedule_race_free_and_in_reference_order(MT, ns):
    seq = ops(2, MT, 128, 0, ns)
    launches = happens_before(seq)
    streams = {o["stream"] for o, _, _, _ in launches}
    assert streams <= {10 + i for i in range(ns)} | {20 + i for i in range(ns)}
    assert check_races(launches) > 0
    got = [(qr_label(o), o["wr"], o["rd"]) for o, _, _, _ in launches]
    hist_g, seen_g = writer_history(got)
    hist_r, seen_r = writer_history(qr_reference(MT))
    assert hist_g == hist_r
    assert seen_g == seen_r


def test_random_schedules_race_free():
    """Random tile counts and tile sizes across every schedule and tier: the
    same race and completeness checks (hypothesis-driven, bounded)."""
    from hypothesis import given, settings, strategies as st

    @settings(max_examples=40, deadline=None, derandomize=True)
    @given(algo=st.sampled_from([0, 1]), MT=st.integers(2, 11), b=st.sampled_from([128, 256, 512]),
           pv=st.sampled_from([(0, 0), (0, 2), (0, 4), (0, 6), (1, 0), (2, 0), (2, 1), (3, 0), (4, 0), (4, 1)]))
    def check(algo, MT, b, pv):
        prec, variant = pv
        seq = ops(algo, MT, b, prec, variant)
        check_races(happens_before(seq))
        check_semantics(seq, algo, MT)

    check()


def test_random_qr_schedules():
    from hypothesis import given, settings, strategies as st

    @settings(max_examples=20, deadline=None, derandomize=True)
    @given(MT=st.integers(2, 10), ns=st.integers(1, 8))
    def check(MT, ns):
        seq = ops(2, MT, 128, 0, ns)
        launches = happens_before(seq)
        check_races(launches)
        got = [(qr_label(o), o["wr"], o["rd"]) for o, _, _, _ in launches]
        assert writer_history(got) == writer_history(qr_reference(MT))

    check()
