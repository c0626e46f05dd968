"""Warm solves or factorizations of one workload for a kernel profiler (no
reference counterpart).

    python -m palial_amd.profile_solve --algo cholesky --precision 2 --n 32768 --tile 512 --kappa 1e2
    python -m palial_amd.profile_solve --factor qr --n 16384 --tile 512

Solves: A = Q diag(s) Q^T with Q from the library's own Householder QR of a
seeded Gaussian (CondGenerator) and a seeded standard normal right-hand side,
both resident on the device.  Factorizations (--factor): the seeded S + N I
(Cholesky), G + N I (LU) or G (QR) of the library's generator, regenerated
before each run.  Everything is launched stream-ordered (no hipGraph replay),
so a kernel tracer sees each dispatch.  One line per run: index, return code,
factorization and solve/refinement milliseconds, refinement steps, backward
error (solves) or TFLOP/s (factorizations).
"""
import argparse
import ctypes as C
import sys

import numpy as np

from . import (CondGenerator, DeviceMatrix, GEN_DENSE, GEN_GAUSS, GEN_SYM, _check, factor_device, finalize,
               init, lib, solve)


def factorizations(a):
    N, b = a.n, a.tile
    A = DeviceMatrix(N, b)
    S = DeviceMatrix(N, b) if a.factor == "qr" else None
    kind, shift = {"cholesky": (GEN_SYM, N), "lu": (GEN_DENSE, N), "qr": (GEN_GAUSS, 0.0)}[a.factor]
    try:
        for i in range(a.solves):
            A.generate(kind, seed=a.seed, shift=shift)
            rc, rep = factor_device(a.factor, A, S, no_graph=True)
            print("factor", i, rc, round(rep.factor_ms, 3), round(rep.tflops, 2), flush=True)
    finally:
        A.free()
        if S is not None:
            S.free()
        finalize()
    print("done", flush=True)
    return 0


def variants(a, A, dX, rhs):
    """Each variant q:pair -- solves with sweep block q and paired updates
    on or off -- timed on the host around the synchronous call (A and b in
    HBM), the first solve untimed; one JSON line per variant."""
    import json
    import time
    from . import set_option, synchronize
    N, b = a.n, a.tile
    prev = set_option("pair", 0)
    rc_all = 0
    try:
        for item in a.variants.split(","):
            q, pr = (int(x) for x in item.split(":"))
            set_option("pair", pr)
            rows = []
            for i in range(a.solves + 1):
                _check(lib().palial_hip_memcpy_h2d(dX, rhs.ctypes.data, N * a.nrhs * 8))
                synchronize()
                t0 = time.perf_counter()
                _, rep, rc = solve(A, (dX, a.nrhs, N), algo=a.algo, precision=a.precision, device=True,
                                   factor_tile=b, check=False, sweep_block=q)
                ms = (time.perf_counter() - t0) * 1e3
                if i:
                    rows.append((ms, rep.factor_ms, rep.solve_ms, rep.iterations, rep.backward_error, rc,
                                 rep.precision_used))
            ms = sorted(r[0] for r in rows)
            out = {"sweep_block": q, "pair": pr, "ms_median": round(ms[len(ms) // 2], 3),
                   "factor_ms": round(rows[-1][1], 3), "solve_ms": round(rows[-1][2], 3),
                   "iterations": rows[-1][3], "backward_error": rows[-1][4], "rc": rows[-1][5],
                   "precision_used": rows[-1][6]}
            rc_all = rc_all or rows[-1][5]
            print(json.dumps(out), flush=True)
    finally:
        set_option("pair", prev)
    return 0 if rc_all == 0 else 1


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--algo", choices=["cholesky", "lu"], default="cholesky")
    ap.add_argument("--precision", type=int, default=2)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--tile", type=int, default=512, help="tile of the matrix and of the factorization")
    ap.add_argument("--kappa", type=float, default=1e2)
    ap.add_argument("--nrhs", type=int, default=1)
    ap.add_argument("--solves", type=int, default=3)
    ap.add_argument("--sweep-block", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--factor", choices=["cholesky", "lu", "qr"], default=None,
                    help="time FP64 factorizations of the library's seeded inputs instead of solves")
    ap.add_argument("--variants", default=None,
                    help="q:pair,... -- for each sweep block q and paired-update setting, the solves "
                         "(the first untimed), one JSON line each")
    a = ap.parse_args(argv)
    init(0)
    if a.factor:
        return factorizations(a)
    N, b = a.n, a.tile
    gen = CondGenerator(N, b, seed=a.seed)
    A = DeviceMatrix(N, b)
    gen.fill(A, a.kappa)
    gen.free()
    rhs = np.asfortranarray(np.random.default_rng(a.seed + 1).standard_normal((N, a.nrhs)))
    dX = C.c_void_p()
    _check(lib().palial_hip_malloc(C.byref(dX), N * a.nrhs * 8))
    rc_all = 0
    if a.variants:
        try:
            return variants(a, A, dX, rhs)
        finally:
            lib().palial_hip_free(dX)
            A.free()
            finalize()
    try:
        for i in range(a.solves):
            _check(lib().palial_hip_memcpy_h2d(dX, rhs.ctypes.data, N * a.nrhs * 8))
            _, rep, rc = solve(A, (dX, a.nrhs, N), algo=a.algo, precision=a.precision, device=True,
                               factor_tile=b, check=False, sweep_block=a.sweep_block)
            print("solve", i, rc, round(rep.factor_ms, 3), round(rep.solve_ms, 3), rep.iterations,
                  "%.3g" % rep.backward_error, flush=True)
            rc_all = rc_all or rc
    finally:
        lib().palial_hip_free(dX)
        A.free()
        finalize()
    print("done", flush=True)
    return 0 if rc_all == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
