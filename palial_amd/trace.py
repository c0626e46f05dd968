"""Per-task trace lines in the reference tracer's format for the Cholesky, LU
and QR factorizations: one `name, ms` line per tile task, in task-creation order,
ms with six decimals (palial_trace/log.h:43-50).

Names are built as src/cholesky.h:14-173 builds them: the task kind followed
by k, m and n printed back to back, where m and n are whatever the loop
variables hold when the task is created (so potrf of step k >= 1 carries the
m = MT, n = MT - 1 that the previous step's loops left behind, and syrk
carries the n of the previous column's gemm loop).  LU names follow
src/lu.h:13-173 and QR names src/qr.h:17-185 the same way; their k = 0 tasks
print loop variables the reference never initialises (src/lu.h:13,
src/qr.h:17), printed here as 0.  A QR update's three GEMM phases are split
evenly over its n > k tiles.
The Cholesky names are pinned to the reference's recorded trace; the
reference holds no recorded LU or QR trace.

Times come from the library's per-launch records (profile(2) before a
factorization or solve, then profile_records()).  A launch of the chain
schedule covers one kind of tile task of one step; its time is split over its
tasks in proportion to their flops (potrf b^3/3, trsm and syrk b^3, gemm
2 b^3).  Packing launches have no reference task and are left out, so the
lines add up to the launches' time without them.  Launches are matched to the
schedule's enumeration (costmodel.launches) group by group: within one
(class, flops, bytes) shape both lists run in step order."""
import numpy as np

from . import costmodel as cm


def cholesky_tasks(MT):
    """[(name, task)] in creation order; task = ("potrf", k), ("trsm", k, m),
    ("syrk", k, m) or ("gemm", k, m, n) (gemm updates tile (n, m))."""
    out = []
    k = m = n = 0
    for k in range(MT):
        out.append((f"potrf{k}{m}{n}", ("potrf", k)))
        m = k + 1
        while m < MT:
            out.append((f"trsm{k}{m}{n}", ("trsm", k, m)))
            m += 1
        m = k + 1
        while m < MT:
            out.append((f"syrk{k}{m}{n}", ("syrk", k, m)))
            n = k + 1
            while n < m:
                out.append((f"gemm{k}{m}{n}", ("gemm", k, m, n)))
                n += 1
            m += 1
    return out


def lu_tasks(MT):
    """[(name, task)] in creation order; task = ("getrf", k), ("trsmL", k, m)
    (tile (m, k)), ("trsmU", k, n) (tile (k, n)) or ("gemm", k, m, n) (tile (m, n))."""
    out = []
    k = m = n = 0
    for k in range(MT):
        out.append((f"getrf{k}{m}{n}", ("getrf", k)))
        m = k + 1
        while m < MT:
            out.append((f"trsm{k}{m}{n}", ("trsmL", k, m)))
            m += 1
        n = k + 1
        while n < MT:
            out.append((f"trsm{k}{m}{n}", ("trsmU", k, n)))
            m = k + 1
            while m < MT:
                out.append((f"gemm{k}{m}{n}", ("gemm", k, m, n)))
                m += 1
            n += 1
    return out


def _covered_lu(k, stream, cls, MT, split, idx):
    """Reference tasks of one enumerated launch of the LU chain schedule; idx:
    0 for the first launch of its (step, stream, class), 1 for the second (the
    row panel's TRSM is issued before the column panel's)."""
    if cls == cm.CLASSES["diag"]:
        return [("getrf", k)]
    if cls == cm.CLASSES["trsm"]:
        tiles = [k + 1] if stream == "chain" else list(range(k + 2, MT))
        return [("trsmU", k, n) for n in tiles] if idx == 0 else [("trsmL", k, m) for m in tiles]
    if cls == cm.CLASSES["pack"]:
        return []
    if stream == "chain":                              # tile (k+1, k+1)
        return [("gemm", k, k + 1, k + 1)]
    if not split:
        return [("gemm", k, m, n) for n in range(k + 1, MT) for m in range(k + 1, MT) if (m, n) != (k + 1, k + 1)]
    if stream == "first":                              # row and column k+1, and tile (k+2, k+2)
        return [("gemm", k, k + 1, n) for n in range(k + 2, MT)] + \
               [("gemm", k, m, k + 1) for m in range(k + 2, MT)] + [("gemm", k, k + 2, k + 2)]
    return [("gemm", k, m, n) for n in range(k + 2, MT) for m in range(k + 2, MT) if (m, n) != (k + 2, k + 2)]


def qr_tasks(MT):
    """[(name, task)] in creation order (src/qr.h:17-185): ("geqrt", k),
    ("ormqr", k, n), ("tsqrt", k, m), ("tsmqr", k, m, n)."""
    out = []
    k = m = n = 0
    for k in range(MT):
        out.append((f"geqrt{k}{m}{n}", ("geqrt", k)))
        n = k + 1
        while n < MT:
            out.append((f"ormqr{k}{m}{n}", ("ormqr", k, n)))
            n += 1
        m = k + 1
        while m < MT:
            out.append((f"tsqrt{k}{m}{n}", ("tsqrt", k, m)))
            n = k + 1
            while n < MT:
                out.append((f"tsmqr{k}{m}{n}", ("tsmqr", k, m, n)))
                n += 1
            m += 1
    return out


def qr_launches(N, b):
    """The records of one QR factorization in issue order (dag_qr,
    palial_amd/csrc/api.hip): (group, class, flops, bytes); group is a task
    for geqrt and tsqrt, ("ormqr", k) or ("tsmqr", k, m) for the three GEMM
    phases of an update (W = V^T C, W = T^T W, C -= V W over every n > k)."""
    MT = N // b
    b3, tb = float(b) ** 3, 8.0 * b * b
    D, Q, U = cm.CLASSES["diag"], cm.CLASSES["tsqrt"], cm.CLASSES["qrupd"]
    out = []
    for k in range(MT):
        out.append((("geqrt", k), D, 4 * b3 / 3, 3 * tb))
        T = MT - k - 1
        if T == 0:
            continue
        out += [(("ormqr", k), U, T * b3, 3 * T * tb)] * 3
        for m in range(k + 1, MT):
            out.append((("tsqrt", k, m), Q, 2 * b3, 5 * tb))
            out += [(("tsmqr", k, m), U, 2 * T * b3, 4 * T * tb), (("tsmqr", k, m), U, T * b3, 4 * T * tb),
                    (("tsmqr", k, m), U, 2 * T * b3, 3 * T * tb)]
    return out


def qr_task_times(records, N, b):
    """{task: ms} for one QR factorization: geqrt and tsqrt records are tasks;
    an update's three phase records are split evenly over its n > k tiles."""
    MT = N // b
    enum = qr_launches(N, b)
    ms_of = _match(records, [(c, f, by) for _, c, f, by in enum], f"qr schedule of N={N}, b={b}")
    times = {}
    for (grp, _, _, _), ms in zip(enum, ms_of):
        if grp[0] in ("geqrt", "tsqrt"):
            tasks = [grp]
        else:
            tasks = [grp[:2] + (n,) if grp[0] == "ormqr" else grp + (n,) for n in range(grp[1] + 1, MT)]
        for t in tasks:
            times[t] = times.get(t, 0.0) + ms / len(tasks)
    return times


def _match(records, enum, what):
    """Recorded ms per enumerated (class, flops, bytes), matched group by group in order."""
    rec = np.asarray(records, dtype=np.float64).reshape(-1, 6)
    key = lambda c, f, by: (int(c), int(round(f)), int(round(by)))
    groups = {}
    for i, (c, f, by) in enumerate(enum):
        groups.setdefault(key(c, f, by), []).append(i)
    seen = {g: 0 for g in groups}
    ms_of = [None] * len(enum)
    for c, _, _, f, by, ms in rec:
        g = key(c, f, by)
        if g not in groups or seen[g] >= len(groups[g]):
            raise ValueError(f"record {g} has no launch left in the {what}")
        ms_of[groups[g][seen[g]]] = ms
        seen[g] += 1
    if any(t is None for t in ms_of):
        raise ValueError("records cover only part of the schedule: pass the records of one factorization")
    return ms_of


def _covered(k, stream, cls, MT, split):
    """Reference tasks of one enumerated launch of the Cholesky chain schedule."""
    if cls == cm.CLASSES["diag"]:
        return [("potrf", k)]
    if cls == cm.CLASSES["trsm"]:
        return [("trsm", k, k + 1)] if stream == "chain" else [("trsm", k, m) for m in range(k + 2, MT)]
    if cls == cm.CLASSES["pack"]:
        return []
    if stream == "chain":                              # tile (k+1, k+1)
        return [("syrk", k, k + 1)]
    if not split:                                      # the rest of step k's update in one launch
        return [("syrk", k, m) for m in range(k + 2, MT)] + \
               [("gemm", k, m, n) for m in range(k + 2, MT) for n in range(k + 1, m)]
    if stream == "first":                              # tile row k+1 and tile (k+2, k+2)
        return [("gemm", k, m, k + 1) for m in range(k + 2, MT)] + [("syrk", k, k + 2)]
    return [("syrk", k, m) for m in range(k + 3, MT)] + \
           [("gemm", k, m, n) for m in range(k + 3, MT) for n in range(k + 2, m)]


def _flops(task, b):
    b3 = float(b) ** 3
    return {"potrf": b3 / 3, "getrf": 2 * b3 / 3, "trsm": b3, "trsmL": b3, "trsmU": b3, "syrk": b3,
            "gemm": 2 * b3}[task[0]]


def task_times(algo, records, N, b, prec=0, split=True):
    """{task: ms} from the per-launch records of one factorization (the
    records of exactly one factorization, in issue order)."""
    if algo == "qr":
        return qr_task_times(records, N, b)
    MT = N // b
    enum = cm.launches(algo, N, b, prec, split=split)
    ms_of = _match(records, [(c, f, by) for _, _, c, f, by in enum], f"{algo} schedule of N={N}, b={b}")
    times = {}
    occ = {}
    for (k, st, c, f, by), ms in zip(enum, ms_of):
        idx = occ.get((k, st, c), 0)
        occ[(k, st, c)] = idx + 1
        tasks = _covered(k, st, c, MT, split) if algo == "cholesky" else _covered_lu(k, st, c, MT, split, idx)
        tot = sum(_flops(t, b) for t in tasks)
        for t in tasks:
            times[t] = times.get(t, 0.0) + ms * _flops(t, b) / tot
    return times


def cholesky_task_times(records, N, b, prec=0, split=True):
    return task_times("cholesky", records, N, b, prec, split)


def trace_lines(algo, records, N, b, prec=0, split=True):
    """The reference tracer's `name, ms` lines for one factorization."""
    times = task_times(algo, records, N, b, prec, split)
    tasks = {"cholesky": cholesky_tasks, "lu": lu_tasks, "qr": qr_tasks}[algo](N // b)
    return [f"{name}, {times[task]:f}" for name, task in tasks]


def cholesky_trace_lines(records, N, b, prec=0, split=True):
    return trace_lines("cholesky", records, N, b, prec, split)


def main(argv=None):
    """python -m palial_amd.trace [-a cholesky|lu|qr] -m N -b B: one profiled
    FP64 factorization of a generated matrix (S + N I; G + N I for LU; G for
    QR) on the GPU, printed as the reference tracer prints its run."""
    import argparse
    from . import init, DeviceMatrix, GEN_SYM, GEN_DENSE, profile, profile_records, factor_device
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("-a", choices=("cholesky", "lu", "qr"), default="cholesky")
    ap.add_argument("-m", type=int, default=8192, help="matrix size N")
    ap.add_argument("-b", type=int, default=512, help="tile size")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    if args.m % args.b:
        ap.error("the tile size must divide N")
    init(0)
    A = DeviceMatrix(args.m, args.b)
    A.generate(GEN_SYM if args.a == "cholesky" else GEN_DENSE, seed=args.seed,
               shift=0.0 if args.a == "qr" else float(args.m))
    S = DeviceMatrix(args.m, args.b) if args.a == "qr" else None
    profile(2)
    info, _ = factor_device(args.a, A, S=S)
    rec = profile_records()
    profile(0)
    A.free()
    if S is not None:
        S.free()
    if info != 0:
        raise SystemExit(f"factorization failed (info {info})")
    try:
        lines = trace_lines(args.a, rec, args.m, args.b)
    except ValueError:                                  # the schedule without the split update
        lines = trace_lines(args.a, rec, args.m, args.b, split=False)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    from . import finalize
    rc = main()
    finalize()
    raise SystemExit(rc)
