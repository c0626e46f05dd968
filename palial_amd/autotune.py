"""Autotuner for the solve path: picks the factorization tile size and the
precision tier with the task-time cost model (palial_amd.costmodel).

    python -m palial_amd.autotune --n 32768 --algo cholesky

1. collect: runs one profiled solve per candidate (tile size, tier) on the
   device and keeps every launch's device time (palial_hip_profile(2)), the
   factorization and refinement times and the iteration count;
2. trains the cost model on the launches (GPU kernels);
3. predicts each candidate's factorization time from its launch list (the
   chain and bulk streams overlap step by step) and adds the refinement time
   measured for that tier; reports the best candidate.
The exhaustive search runs over b in {256, 512} (those dividing N) and the
FP64 / FP32 / BF16 tiers.  Stream priorities are not searched: every stream
runs at the default priority."""
import argparse
import json
import sys

import numpy as np

from . import costmodel as cm


def candidates(N, tiles=(256, 512), tiers=(0, 1, 2)):
    return [(b, t) for b in tiles if N % b == 0 for t in tiers]


def collect(P, A, rhs_ptr, rhs, N, algo, cands):
    """Profiled solves: returns (records (n, 6), per-candidate measurements)."""
    L = P.lib()
    recs, meas = [], {}
    for b, tier in cands:
        P._check(L.palial_hip_memcpy_h2d(rhs_ptr, rhs.ctypes.data, N * 8))
        P.synchronize()
        P.profile(2)
        _, rep, rc = P.solve(A, (rhs_ptr, 1, N), algo=algo, precision=tier, device=True, max_iter=40,
                             check=False, factor_tile=b)
        r = P.profile_records()
        P.profile(0)
        recs.append(r)
        meas[(b, tier)] = {"factor_ms": rep.factor_ms, "refine_ms": rep.solve_ms, "iterations": rep.iterations,
                           "converged": rc == 0, "launches": len(r)}
    return np.concatenate(recs), meas


def tune(P, A, N, algo="cholesky", tiles=(256, 512), tiers=(0, 1, 2), epochs=300, seed=1):
    import ctypes as C
    cands = candidates(N, tiles, tiers)
    rhs = np.random.default_rng(seed).standard_normal(N)
    dX = C.c_void_p()
    P._check(P.lib().palial_hip_malloc(C.byref(dX), N * 8))
    try:
        recs, meas = collect(P, A, dX, rhs, N, algo, cands)
    finally:
        P.lib().palial_hip_free(dX)
    X, y = cm.featurize(recs)
    model = cm.MLP(seed=seed)
    loss = model.train(X, y, epochs=epochs)
    scored = []
    for b, tier in cands:
        m = meas[(b, tier)]
        pred = cm.predict_factor_ms(model, algo, N, b, tier)
        scored.append({"b": b, "tier": tier, "predicted_factor_ms": pred, "measured_factor_ms": m["factor_ms"],
                       "refine_ms": m["refine_ms"], "iterations": m["iterations"], "converged": m["converged"],
                       "predicted_total_ms": pred + m["refine_ms"]})
    ok = [s for s in scored if s["converged"]]
    best = min(ok, key=lambda s: s["predicted_total_ms"]) if ok else None
    return {"best": best, "candidates": scored, "train_loss": [float(loss[0]), float(loss[-1])],
            "samples": int(len(X))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--algo", default="cholesky")
    ap.add_argument("--kappa", type=float, default=1e2)
    args = ap.parse_args()
    import palial_amd as P
    P.init(0)
    A = P.DeviceMatrix(args.n, 256)
    A1, A2 = P.kron_spd_factors(args.n, args.kappa, seed=0)
    A.generate_kron(A1, A2)
    res = tune(P, A, args.n, args.algo)
    A.free()
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
