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
src/qr.h:17), printed here as 0.  The Cholesky names are pinned to the
reference's recorded trace; the reference holds no recorded LU or QR trace.
The library builds the same list (palial_hip_task_name) for the reference's
trace hooks.

Times are measured on the device per task (palial_hip_trace): a task inside a
batched launch from the timestamps of its first and last workgroup, a
diagonal-tile or QR panel task from HIP events around its launch; a QR
update's time is the sum over its three launches (W = V^T C, W = T^T W,
C -= V W).  `dag_edges` gives the task DAG the reference's tracer records from
OpenMP dependences (palial_trace/palial_trace.c:131-202), derived here from
the tiles each task reads and writes (the reference's LU names wrong tiles in
two depend clauses, src/lu.h:62,104; the edges here are the correct ones)."""
import numpy as np

# library task kinds (palial_hip_task_name) <-> the tuples of this module
KINDS = {0: "potrf", 1: "trsm", 2: "syrk", 3: "gemm", 4: "getrf", 5: "trsmL", 6: "trsmU", 7: "gemm",
         8: "geqrt", 9: "ormqr", 10: "tsqrt", 11: "tsmqr"}


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


def tasks(algo, MT):
    return {"cholesky": cholesky_tasks, "lu": lu_tasks, "qr": qr_tasks}[algo](MT)


def _key(task):
    """(kind code, k, m, n) of a task tuple, as the library numbers them."""
    kind = task[0]
    if kind in ("potrf", "getrf", "geqrt"):
        return ({"potrf": 0, "getrf": 4, "geqrt": 8}[kind], task[1], -1, -1)
    if kind in ("trsm", "syrk", "trsmL", "tsqrt"):
        return ({"trsm": 1, "syrk": 2, "trsmL": 5, "tsqrt": 10}[kind], task[1], task[2], -1)
    if kind in ("trsmU", "ormqr"):
        return ({"trsmU": 6, "ormqr": 9}[kind], task[1], -1, task[2])
    return (None, task[1], task[2], task[3])      # gemm / tsmqr: the code depends on the algorithm


def trace_lines(algo, records, N, b):
    """The reference tracer's `name, ms` lines from the library's per-task
    records of one factorization (palial_amd.trace_records())."""
    rec = np.asarray(records, dtype=np.float64).reshape(-1, 6)
    tl = tasks(algo, N // b)
    if len(rec) != len(tl):
        raise ValueError(f"{len(rec)} records for {len(tl)} {algo} tasks of N={N}, b={b}")
    gemm_code = {"cholesky": 3, "lu": 7, "qr": 11}[algo]
    out = []
    for (name, task), r in zip(tl, rec):
        code, k, m, n = _key(task)
        code = gemm_code if code is None else code
        if (code, k, m, n) != tuple(int(x) for x in r[:4]):
            raise ValueError(f"record {r[:4]} is not task {name}")
        out.append(f"{name}, {r[4]:f}")
    return out


def tiles_of(algo, task):
    """(read tiles, written tiles) of a reference task, tiles as (row, col)."""
    kind, k = task[0], task[1]
    if algo == "cholesky":
        if kind == "potrf":
            return [], [(k, k)]
        if kind == "trsm":
            return [(k, k)], [(k, task[2])]
        if kind == "syrk":
            return [(k, task[2])], [(task[2], task[2])]
        m, n = task[2], task[3]                       # A(n,m) -= A(k,n)^T A(k,m)
        return [(k, n), (k, m)], [(n, m)]
    if algo == "lu":
        if kind == "getrf":
            return [], [(k, k)]
        if kind == "trsmL":
            return [(k, k)], [(task[2], k)]
        if kind == "trsmU":
            return [(k, k)], [(k, task[2])]
        m, n = task[2], task[3]                       # A(m,n) -= A(m,k) A(k,n)
        return [(m, k), (k, n)], [(m, n)]
    # QR: V / T tiles are read, R / A tiles written; S(m,k) carries T with A(m,k)
    if kind == "geqrt":
        return [], [(k, k)]
    if kind == "ormqr":
        return [(k, k)], [(k, task[2])]
    if kind == "tsqrt":
        return [], [(k, k), (task[2], k)]
    m, n = task[2], task[3]                           # rows k and m of column n, V = A(m,k)
    return [(m, k)], [(k, n), (m, n)]


def dag_edges(algo, MT):
    """Dependences of the task DAG in the sequential semantics: for every task
    (creation index i) the earlier tasks it must follow -- the last writer of
    each tile it reads or writes, and the readers of a tile since its last
    write before it writes that tile.  Returns {i: sorted list of j < i}."""
    tl = tasks(algo, MT)
    last_w, readers = {}, {}
    edges = {}
    for i, (_, t) in enumerate(tl):
        rd, wr = tiles_of(algo, t)
        dep = set()
        for x in rd + wr:
            if x in last_w:
                dep.add(last_w[x])
        for x in wr:
            dep.update(readers.get(x, ()))
        # a QR task reading A(m,k) written by its own step's tsqrt is covered
        # above; drop self-references
        dep.discard(i)
        edges[i] = sorted(dep)
        for x in rd:
            readers.setdefault(x, set()).add(i)
        for x in wr:
            last_w[x] = i
            readers[x] = set()
    return edges


def measure(algo, N, b, seed=0, precision=0):
    """One traced factorization on the GPU of the generated input (S + N I for
    Cholesky, G + N I for LU, G for QR; precision > 0: a Cholesky/LU solve in
    that tier, whose factorization is traced).  Returns the library's records."""
    from . import (init, DeviceMatrix, GEN_SYM, GEN_DENSE, factor_device, solve, trace_enable,
                   trace_records)
    init(0)
    A = DeviceMatrix(N, b)
    S = DeviceMatrix(N, b) if algo == "qr" else None
    try:
        A.generate(GEN_SYM if algo == "cholesky" else GEN_DENSE, seed=seed,
                   shift=0.0 if algo == "qr" else float(N))
        trace_enable(True)
        if precision == 0:
            info, _ = factor_device(algo, A, S=S)
            if info != 0:
                raise RuntimeError(f"factorization failed (info {info})")
        else:
            import ctypes as C
            from . import lib, _check
            rhs = np.random.default_rng(seed).standard_normal(N)
            dX = C.c_void_p()
            _check(lib().palial_hip_malloc(C.byref(dX), N * 8))
            _check(lib().palial_hip_memcpy_h2d(dX, rhs.ctypes.data, N * 8))
            _, rep, rc = solve(A, (dX, 1, N), algo=algo, precision=precision, device=True, factor_tile=b,
                               check=False)
            lib().palial_hip_free(dX)
            if rc != 0:
                raise RuntimeError(f"solve failed ({rc})")
        return trace_records()
    finally:
        trace_enable(False)
        A.free()
        if S is not None:
            S.free()


def main(argv=None):
    """python -m palial_amd.trace [-a cholesky|lu|qr] -m N -b B [--edges]: one
    traced FP64 factorization of a generated matrix (S + N I; G + N I for LU;
    G for QR) on the GPU, printed as the reference tracer prints its run;
    --edges prints each task's dependences instead."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("-a", choices=("cholesky", "lu", "qr"), default="cholesky")
    ap.add_argument("-m", type=int, default=8192, help="matrix size N")
    ap.add_argument("-b", type=int, default=512, help="tile size")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", type=int, default=0, help="tier of a traced solve (0: plain FP64 factorization)")
    ap.add_argument("--edges", action="store_true", help="print the task DAG's dependences")
    args = ap.parse_args(argv)
    if args.m % args.b:
        ap.error("the tile size must divide N")
    if args.edges:
        tl = tasks(args.a, args.m // args.b)
        for i, deps in dag_edges(args.a, args.m // args.b).items():
            print(f"{tl[i][0]}: " + " ".join(tl[j][0] for j in deps))
        return 0
    rec = measure(args.a, args.m, args.b, seed=args.seed, precision=args.precision)
    print("\n".join(trace_lines(args.a, rec, args.m, args.b)))
    return 0


if __name__ == "__main__":
    from . import finalize
    rc = main()
    finalize()
    raise SystemExit(rc)
