"""Random configurations of the library's API on the CPU build of its HIP
sources (tests/emu): factorizations of every algorithm and schedule, and
solves of every precision tier, sweep block and paired-update setting, on
small matrices with tile sizes from 32 to 256 (solves: multiples of 64, as
palial_hip_solve requires).  Each result is checked -- factors
against the oracle to 1e-10, solves to backward error 2e-14 recomputed with
numpy, or a refusal with PALIAL_ERR_ARG -- and, on the race-checking variant,
every access against the stream order.

    python tests/emu/fuzz.py SEED COUNT        (FUZZ_VARIANT=race|asan|plain)

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


def main(seed, count):
    P.init(0)
    L = P.lib()
    if VARIANT == "race":
        L.hipemu_races.restype = C.c_long
    rng = random.Random(seed)
    fails = 0
    for it in range(count):
        kind = rng.choice(["factorization", "factorization", "solve", "solve", "solve"])
        # solves take tile sizes b % 64 == 0 (palial_hip_solve); factorizations any
        b = rng.choice([32, 64, 96, 128, 128, 256] if kind == "factorization" else [64, 128, 128, 256])
        MT = rng.randint(1, 7 if b <= 128 else 4)
        N = b * MT
        t0 = time.time()
        try:
            ok, desc = factorization(rng, it, N, b) if kind == "factorization" else solve(rng, it, N, b, MT)
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
