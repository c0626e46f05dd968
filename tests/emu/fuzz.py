"""Random configurations of the library's API on the CPU build of its HIP
sources (tests/emu): factorizations of every algorithm and schedule, solves
of every precision tier, sweep block and paired-update setting (also on
matrices of condition number up to 1e9, through breakdown, shifted
refactorization, GMRES-IR and the FP64 fallback), batched factorizations,
and QR least-squares solves and Q formation, on
small matrices with tile sizes from 32 to 256 (solves: multiples of 64, as
palial_hip_solve requires).  Each result is checked -- factors against the
oracle to 1e-10, solves to backward error 2e-14 recomputed with numpy (or a
refusal with PALIAL_ERR_ARG), batched factors against numpy Cholesky and a
Doolittle elimination, QR by its residuals and numpy's solution -- and, on
the race-checking variant, every access against the stream order.

    python tests/emu/fuzz.py SEED COUNT        (FUZZ_VARIANT=race|asan|plain)

FUZZ_ROWS=LO-HI draws the tile rows of factorizations and solves from LO..HI
instead (with tile sizes 32 and 64 for factorizations, 64 for solves), for
deeper schedules.

One line per configuration ("ok" or "BAD"), then "fails <n>".
"""
import ctypes as C
import os
import random
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import palial_amd as P  # noqa: E402
from emu import build as emu_build  # noqa: E402

VARIANT = os.environ.get("FUZZ_VARIANT", "race")
P.LIB_PATH = emu_build.build(VARIANT)
import oracle  # noqa: E402

ERR_ARG = -1
ROWS = tuple(int(x) for x in os.environ["FUZZ_ROWS"].split("-")) if os.environ.get("FUZZ_ROWS") else None


def races(L):
    """(count, report) of the race checker since the last call (0 elsewhere)."""
    if VARIANT != "race":
        return 0, ""
    n = L.hipemu_races()
    buf = C.create_string_buffer(1 << 14)
    L.hipemu_race_report(buf, len(buf))
    L.hipemu_race_reset()
    return n, buf.value.decode()


def backward_error(A, X, B):
    R = B - A @ X
    return max(np.abs(R[:, j]).max() / (np.abs(A).sum(1).max() * np.abs(X[:, j]).max() + np.abs(B[:, j]).max())
               for j in range(B.shape[1]))


def factorization(rng, it, N, b):
    algo = rng.choice(["cholesky", "lu", "qr"])
    kw = {"schedule": rng.choice([0, 1, 2, 3, 8] if algo == "qr" else [0, 1, 2, 3, 4]),
          "no_graph": rng.random() < 0.7}
    a = oracle.gen("sym" if algo == "cholesky" else "dense", N, b, seed=it, shift=0.0 if algo == "qr" else N)
    x = a.copy()
    if algo == "qr":
        info, _ = P.qr(x, np.zeros(N * N), b=b, **kw)
        ref, _ = oracle.qr(a, N, b)
    elif algo == "lu":
        info, _, _ = P.lu(x, b=b, **kw)
        ref = oracle.lu(a, N, b)
    else:
        info, _ = P.cholesky(x, b=b, **kw)
        ref, _ = oracle.cholesky(a, N, b)
    err = np.abs(x - ref).max() / np.abs(ref).max()
    return info == 0 and err < 1e-10, f"{algo} N={N} b={b} {kw} err={err:.1e} info={info}"


def solve(rng, it, N, b, MT):
    algo = rng.choice(["cholesky", "lu"])
    prec = rng.choice([P.PREC_FP64, P.PREC_FP32, P.PREC_BF16, P.PREC_FP8, P.PREC_BF16X3])
    if prec in (P.PREC_FP8, P.PREC_BF16X3) and (b % 128 or MT < 2):
        prec = P.PREC_BF16           # those tiers need packed panels of a factor tile below N
    nrhs = rng.choice([1, 1, 2, 3, 5, 16])
    q = rng.choice([0, 1, 1, 2, 4])
    pair = rng.random() < 0.3
    a = oracle.gen("sym" if algo == "cholesky" else "dense", N, b, seed=it, shift=N)
    A = oracle.to_dense(a, N, b)
    B = np.random.default_rng(it).standard_normal((N, nrhs))
    P.set_option("pair", int(pair))
    try:
        X, rep, rc = P.solve(a, B, algo=algo, precision=prec, b=b, factor_tile=b, sweep_block=q, check=False)
    finally:
        P.set_option("pair", 0)
    eta = backward_error(A, X, B) if rc == 0 else float("nan")
    # a sweep block the solve cannot take is refused; anything else must converge
    ok = (rc == 0 and eta <= 2e-14) or (rc == ERR_ARG and q > 1)
    return ok, (f"{algo} prec={prec} N={N} b={b} nrhs={nrhs} q={q} pair={pair} rc={rc} eta={eta:.1e} "
                f"used={rep.precision_used}")


def conditioned(rng, it):
    """A solve of Q diag(s) Q^T with condition number 1e2 to 1e9 in a random
    tier: a low-precision factor may break down, be shifted, refine with
    GMRES-IR or hand over to FP64 -- the answer must still reach 2e-14."""
    b = 128
    N = b * rng.randint(2, 5)
    kappa = rng.choice([1e2, 1e5, 1e7, 1e9])
    algo = rng.choice(["cholesky", "lu"])
    prec = rng.choice([P.PREC_FP32, P.PREC_BF16, P.PREC_FP8, P.PREC_BF16X3])
    g = np.random.default_rng(it)
    Q, _ = np.linalg.qr(g.standard_normal((N, N)))
    A = (Q * np.geomspace(1.0, 1.0 / kappa, N)) @ Q.T
    A = (A + A.T) / 2
    a = oracle.from_dense(A, N, b)
    B = g.standard_normal((N, rng.choice([1, 3])))
    X, rep, rc = P.solve(a, B, algo=algo, precision=prec, b=b, factor_tile=b, check=False)
    eta = backward_error(A, X, B) if rc == 0 else float("nan")
    return rc == 0 and eta <= 2e-14, (f"{algo} prec={prec} N={N} kappa={kappa:g} nrhs={B.shape[1]} rc={rc} "
                                      f"eta={eta:.1e} used={rep.precision_used} shifts={rep.shifts} "
                                      f"gmres={rep.gmres_iterations}")


def doolittle(a):
    """No-pivot LU (L unit lower strictly below, U upper), in a copy."""
    a = a.copy()
    for k in range(a.shape[0] - 1):
        a[k + 1:, k] /= a[k, k]
        a[k + 1:, k + 1:] -= np.outer(a[k + 1:, k], a[k, k + 1:])
    return a


def batched(rng, it):
    count = rng.randint(1, 12)
    n = np.array([rng.choice([1, 2, 17, 32, 33, 64, 100, 129, 200]) for _ in range(count)], dtype=np.int32)
    kind = np.array([rng.randint(0, 1) for _ in range(count)], dtype=np.int32)
    prec = rng.choice([P.PREC_FP64, P.PREC_FP32])
    grid, no_graph = rng.choice([0, 1, 3, 8]), rng.random() < 0.5
    bt = P.Batch(n, kind, precision=prec)
    try:
        bt.generate(it)
        mats0, _ = bt.get()
        rhs = [np.random.default_rng(it + i).standard_normal(m) for i, m in enumerate(n)]
        bt.set(rhs=rhs)
        bt.factor(solve=True, no_graph=no_graph, grid=grid)
        mats, xs = bt.get()
    finally:
        bt.free()
    tol = 1e-12 if prec == P.PREC_FP64 else 2e-5
    worst = 0.0
    for a0, f, x, b, k in zip(mats0, mats, xs, rhs, kind):
        a0 = a0.astype(np.float64)
        ref = np.linalg.cholesky(a0).T if k == 0 else doolittle(a0)
        got = np.triu(f.astype(np.float64)) if k == 0 else f.astype(np.float64)
        xe = np.linalg.solve(a0, b)
        worst = max(worst, np.abs(got - ref).max() / np.abs(ref).max(),
                    np.abs(x.astype(np.float64) - xe).max() / (10 * max(1.0, np.abs(xe).max())))
    ok = (bt.info == 0).all() and worst <= tol
    return ok, f"count={count} n={n.tolist()} kinds={kind.tolist()} prec={prec} grid={grid} no_graph={no_graph} " \
               f"err={worst:.1e}"


def qr_uses(rng, it, N, b):
    """QR, then the least-squares solve and Q formed by the NoTrans application."""
    a = oracle.gen("dense", N, b, seed=it, shift=float(N))
    A = oracle.to_dense(a, N, b)
    x, s = a.copy(), np.zeros(N * N)
    P.qr(x, s, b=b, no_graph=rng.random() < 0.5)
    nrhs = rng.choice([1, 2, 4])
    B = np.random.default_rng(it).standard_normal((N, nrhs))
    X, _ = P.qr_solve(x, s, B, b=b)
    err_x = np.abs(X - np.linalg.solve(A, B)).max() / np.abs(X).max()
    Qt = oracle.gen("identity", N, b)
    P.qr_apply(x, s, Qt, trans=P.NOTRANS, b=b, N=N)
    Q = oracle.to_dense(Qt, N, b)
    R = np.triu(oracle.to_dense(x, N, b))
    eps = np.finfo(float).eps
    r1 = np.linalg.norm(A - Q @ R) / (N * eps * np.linalg.norm(A))
    r2 = np.linalg.norm(np.eye(N) - Q.T @ Q) / (N * eps)
    ok = err_x < 1e-10 and r1 < 10 and r2 < 10
    return ok, f"N={N} b={b} nrhs={nrhs} err_x={err_x:.1e} |A-QR|={r1:.2f} |I-QtQ|={r2:.2f}"


def main(seed, count):
    P.init(0)
    L = P.lib()
    if VARIANT == "race":
        L.hipemu_races.restype = C.c_long
    rng = random.Random(seed)
    fails = 0
    for it in range(count):
        kind = rng.choice(["factorization", "factorization", "solve", "solve", "solve", "batched", "qr",
                           "conditioned"])
        # solves take tile sizes b % 64 == 0 (palial_hip_solve, palial_hip_qr_solve); factorizations any
        if ROWS and kind in ("factorization", "solve"):
            b = rng.choice([32, 64]) if kind == "factorization" else 64
            MT = rng.randint(*ROWS)
        else:
            b = rng.choice([32, 64, 96, 128, 128, 256] if kind == "factorization" else [64, 128, 128, 256])
            MT = rng.randint(1, 7 if b <= 128 else 4)
        N = b * MT
        t0 = time.time()
        try:
            if kind == "factorization":
                ok, desc = factorization(rng, it, N, b)
            elif kind == "solve":
                ok, desc = solve(rng, it, N, b, MT)
            elif kind == "batched":
                ok, desc = batched(rng, it)
            elif kind == "conditioned":
                ok, desc = conditioned(rng, it)
            else:
                ok, desc = qr_uses(rng, it, N, b)
        except Exception as ex:      # noqa: BLE001 -- reported as a failure of this configuration
            ok, desc = False, f"N={N} b={b}: {ex!r}"
        n, text = races(L)
        if n:
            ok = False
            desc += f" RACES {n}: {text[:600]}"
        print(f"{'ok ' if ok else 'BAD'} {kind} {desc} ({time.time() - t0:.1f} s)", flush=True)
        fails += not ok
    print("fails", fails, flush=True)
    return 1 if fails else 0


if __name__ == "__main__":
    sys.exit(main(int(sys.argv[1]), int(sys.argv[2])))
