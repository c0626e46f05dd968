"""Autotuner: tile size, precision tier and schedule (task priority /
look-ahead) of the tiled factorizations and solves, ranked by the task-time
cost model (palial_amd.costmodel), with the exhaustive search as fallback and
as the check of the model's predictions.

    python -m palial_amd.autotune --algo cholesky --n 8192 [--solve] [--measure-all]

Candidates:
  factorizations (FP64): b in {128, 256, 384, 512, 768, 1024} dividing N (QR:
      b <= 512) x schedule (Cholesky / LU: 1 diagonal chain + split update,
      2 diagonal chain, 3 row look-ahead, 4 one stream; QR: 1, 2, 4 or 8
      panel chains in flight);
  solves: factor tile b in {128, 256, 384, 512} dividing N x tier (FP64,
      FP32, BF16, FP8, split BF16; FP8 and split BF16 need b % 128 == 0).
Procedure:
  1. measure a subset -- every schedule (or tier) at the default tile and
     every tile at the default schedule (or FP64) -- each once timed and once
     traced (per-task device times, palial_hip_trace);
  2. train the cost model on the traced tasks (the library's HIP kernels) and
     fit its scale on the subset's measured times;
  3. predict every candidate (solves: predicted factorization + the
     refinement time measured for the tier); rank the measured ones by their
     measurement and the others by prediction, measure the pick when it is not
     in the subset, and fall back to measuring every candidate when the
     predictions of the measured candidates (median) or of the pick are off
     by more than `max_error`;
  4. measure_all=True also measures every candidate, to report what the model's
     pick costs against the exhaustive best, the prediction error over all
     candidates and the measurement time the model saved;
  5. solves: for the pick, the sweep block (tiles per correction-sweep step,
     1, 2 or 4 as the tile count allows) and, for BF16 and split BF16, paired
     trailing updates on or off, chosen by measurement.
`tune_batch` picks the persistent launch's workgroup count for a batch of
small factorizations by measuring each candidate (exhaustive)."""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import costmodel as cm

TILES = (128, 256, 384, 512, 768, 1024)
SOLVE_TILES = (128, 256, 384, 512)
TIERS = (0, 1, 2, 3, 4)


def factor_candidates(algo, N, tiles=TILES):
    bs = [b for b in tiles if N % b == 0 and N // b >= 2 and (algo != "qr" or b <= 512)]
    scheds = (1, 2, 4, 8) if algo == "qr" else (1, 2, 3, 4)
    return [(b, s) for b in bs for s in scheds]


def solve_candidates(N, tiles=SOLVE_TILES, tiers=TIERS):
    bs = [b for b in tiles if N % b == 0 and N // b >= 2 and b % 64 == 0]
    return [(b, t) for b in bs for t in tiers if t < 3 or b % 128 == 0]


def subset(cands, default_b, default_second):
    """Every second coordinate at the default tile, every tile at the default
    second coordinate."""
    bs = sorted({b for b, _ in cands})
    db = default_b if default_b in bs else max(b for b in bs if b <= 512)
    return [c for c in cands if c[0] == db or c[1] == default_second]


class FactorBench:
    """Times FP64 factorizations of config 2's inputs (S + N I, G + N I, G)
    from a pristine device copy: device time of the factorization launched
    stream-ordered (palial_opts.no_graph, as the traced runs the model learns
    from), median of the runs after the first."""

    def __init__(self, P, algo, N, seed=0):
        self.P, self.algo, self.N = P, algo, N
        self.mats = {}
        self.seed = seed

    def _mats(self, b):
        if b not in self.mats:
            for v in self.mats.values():
                for m in v:
                    if m is not None:
                        m.free()
            P, N = self.P, self.N
            A0, A = P.DeviceMatrix(N, b), P.DeviceMatrix(N, b)
            S = P.DeviceMatrix(N, b) if self.algo == "qr" else None
            A0.generate(P.GEN_SYM if self.algo == "cholesky" else P.GEN_DENSE, seed=self.seed,
                        shift=0.0 if self.algo == "qr" else float(N))
            self.mats = {b: (A0, A, S)}
        return self.mats[b]

    def run(self, b, sched, reps=4, trace=False):
        P = self.P
        A0, A, S = self._mats(b)
        ms = []
        for _ in range(reps):
            A.copy_from(A0)
            info, rep = P.factor_device(self.algo, A, S=S, schedule=sched, no_graph=True)
            ms.append(rep.factor_ms)
        rec = None
        if trace:
            A.copy_from(A0)
            P.trace_enable(True)
            try:
                P.factor_device(self.algo, A, S=S, schedule=sched)
                rec = P.trace_records()
            finally:
                P.trace_enable(False)
        return float(np.median(ms[1:] if len(ms) > 1 else ms)), rec

    def free(self):
        for v in self.mats.values():
            for m in v:
                if m is not None:
                    m.free()
        self.mats = {}


class SolveBench:
    """Times solves of one device matrix (its tile is the caller's; the
    factorization tile is the candidate's) to backward error 1e-14."""

    def __init__(self, P, A, algo, seed=1):
        self.P, self.A, self.algo = P, A, algo
        N = A.N
        self.rhs = np.random.default_rng(seed).standard_normal(N)
        self.dX = C.c_void_p()
        P._check(P.lib().palial_hip_malloc(C.byref(self.dX), N * 8))

    def _solve(self, b, tier, sweep_block=0):
        P, N = self.P, self.A.N
        P._check(P.lib().palial_hip_memcpy_h2d(self.dX, self.rhs.ctypes.data, N * 8))
        P.synchronize()
        t0 = time.perf_counter()
        _, rep, rc = P.solve(self.A, (self.dX, 1, N), algo=self.algo, precision=tier, device=True, max_iter=60,
                             check=False, factor_tile=b, sweep_block=sweep_block)
        return (time.perf_counter() - t0) * 1e3, rep, rc

    def run(self, b, tier, reps=3, trace=False, sweep_block=0):
        ts, rep, rc = [], None, 0
        for _ in range(reps):
            t, rep, rc = self._solve(b, tier, sweep_block)
            ts.append(t)
        rec = None
        if trace:
            self.P.trace_enable(True)
            try:
                self._solve(b, tier)
                rec = self.P.trace_records()
            finally:
                self.P.trace_enable(False)
        info = {"factor_ms": rep.factor_ms, "refine_ms": rep.solve_ms, "iterations": rep.iterations,
                "converged": rc == 0, "tier_used": rep.precision_used}
        return float(np.median(ts[1:] if len(ts) > 1 else ts)), rec, info

    def free(self):
        if self.dX:
            self.P.lib().palial_hip_free(self.dX)
            self.dX = C.c_void_p()


def _train(runs, epochs, seed):
    X, y = cm.dataset(runs, seed=seed)
    model = cm.MLP(seed=seed)
    loss = model.train(X, y, epochs=epochs, lr=3e-3)
    return model, loss, len(X)


def model_error(pred, measured):
    """Median relative error of the calibrated predictions on the measured candidates."""
    return float(np.median([abs(pred[c] / v - 1.0) for c, v in measured.items()]))


def _report(cands, pred, measured, meas_time, pick, fallback, all_time=None, exhaustive=None):
    out = {"candidates": len(cands), "measured": len(measured), "pick": list(pick), "fallback": fallback,
           "predicted_ms": {f"{c[0]}/{c[1]}": round(pred[c], 3) for c in cands},
           "measured_ms": {f"{c[0]}/{c[1]}": round(v, 3) for c, v in measured.items()},
           "measure_s": round(meas_time, 3)}
    if exhaustive is not None:
        best = min(exhaustive, key=exhaustive.get)
        errs = [abs(pred[c] / exhaustive[c] - 1.0) for c in cands]
        out.update({"exhaustive_pick": list(best), "exhaustive_best_ms": round(exhaustive[best], 3),
                    "pick_ms": round(exhaustive[pick], 3),
                    "pick_loss": round(exhaustive[pick] / exhaustive[best] - 1.0, 4),
                    "median_abs_error": round(float(np.median(errs)), 4),
                    "max_abs_error": round(float(np.max(errs)), 4),
                    "exhaustive_measure_s": round(all_time, 3),
                    "measure_time_saved_s": round(all_time - meas_time, 3)})
    return out


def tune_factor(P, algo, N, tiles=TILES, epochs=300, seed=0, max_error=0.25, measure_all=False):
    """Best (b, schedule) of the FP64 factorization; see the module docstring."""
    cands = factor_candidates(algo, N, tiles)
    default_sched = 8 if algo == "qr" else 1
    sub = subset(cands, 512, default_sched)
    fb = FactorBench(P, algo, N)
    try:
        measured, runs, t_meas = {}, [], 0.0
        for b, s in sorted(sub):
            t0 = time.perf_counter()
            ms, rec = fb.run(b, s, trace=True)
            t_meas += time.perf_counter() - t0
            measured[(b, s)] = ms
            runs.append((algo, N, b, 0, s, rec))
        model, loss, nsamp = _train(runs, epochs, seed)
        raw = {c: cm.predict_ms(model, algo, N, c[0], 0, c[1]) for c in cands}
        scale = cm.calibrate([raw[c] for c in sub], [measured[c] for c in sub])
        pred = {c: scale * raw[c] for c in cands}
        # measured candidates rank by their measurement, the rest by prediction
        pick = min(cands, key=lambda c: measured.get(c, pred[c]))
        if pick not in measured:
            t0 = time.perf_counter()
            measured[pick] = fb.run(pick[0], pick[1])[0]
            t_meas += time.perf_counter() - t0
        fallback = model_error(pred, measured) > max_error or abs(pred[pick] / measured[pick] - 1.0) > max_error
        if fallback:
            for c in cands:
                if c not in measured:
                    t0 = time.perf_counter()
                    measured[c] = fb.run(c[0], c[1])[0]
                    t_meas += time.perf_counter() - t0
            pick = min(cands, key=measured.get)
        exhaustive, t_all = None, None
        if measure_all:
            exhaustive, t_all = {}, 0.0
            for c in cands:
                t0 = time.perf_counter()
                exhaustive[c] = fb.run(c[0], c[1])[0]
                t_all += time.perf_counter() - t0
        out = _report(cands, pred, measured, t_meas, pick, fallback, t_all, exhaustive)
        out.update({"algo": algo, "N": N, "train_samples": nsamp, "train_loss": [float(loss[0]), float(loss[-1])],
                    "scale": scale, "second": "schedule"})
        return out
    finally:
        fb.free()


def sweep_blocks(N, b, nrhs=1):
    """Tiles per correction-sweep step a solve with factor tile b can take."""
    MT = N // b
    return [q for q in (1, 2, 4) if MT % q == 0 and MT >= 2 * q and q * b * (1 if nrhs == 1 else 4) <= 4096]


def tune_knobs(P, sb, b, tier, N, reps=3):
    """For the picked (b, tier): every sweep block (and, for BF16 and split
    BF16, paired updates on and off), measured; returns ((sweep_block, pair),
    {choice: ms}).  The library's pairing setting is restored afterwards."""
    pairs = [0, 1] if tier in (2, 4) else [0]
    res = {}
    prev = P.set_option("pair", 0)
    try:
        for pr in pairs:
            P.set_option("pair", pr)
            for q in sweep_blocks(N, b):
                ms, _, inf = sb.run(b, tier, reps=reps, sweep_block=q)
                if inf["converged"] and inf["tier_used"] == tier:
                    res[(q, pr)] = ms
    finally:
        P.set_option("pair", prev)
    return (min(res, key=res.get) if res else (1, 0)), res


def tune_solve(P, A, algo, tiles=SOLVE_TILES, tiers=TIERS, epochs=300, seed=0, max_error=0.25, measure_all=False,
               knobs=True):
    """Best (factor tile, tier) of a solve of device matrix A: predicted time =
    model factorization time + the refinement time measured for the tier.
    knobs: then the sweep block and (BF16, split BF16) paired updates for that
    pick, by measurement."""
    N = A.N
    cands = solve_candidates(N, tiles, tiers)
    sub = subset(cands, 512, 0)
    sb = SolveBench(P, A, algo)
    try:
        measured, info, runs, t_meas = {}, {}, [], 0.0
        for b, t in sorted(sub):
            t0 = time.perf_counter()
            ms, rec, inf = sb.run(b, t, trace=True)
            t_meas += time.perf_counter() - t0
            measured[(b, t)], info[(b, t)] = ms, inf
            if inf["tier_used"] == t:
                runs.append((algo, N, b, t, 1, rec))
        model, loss, nsamp = _train(runs, epochs, seed)
        refine = {}
        for (b, t), inf in info.items():
            refine.setdefault(t, []).append(measured[(b, t)] - inf["factor_ms"])
        refine = {t: float(np.median(v)) for t, v in refine.items()}
        raw = {c: cm.predict_ms(model, algo, N, c[0], c[1], 1) for c in cands}
        fac = [(raw[c], info[c]["factor_ms"]) for c in sub if info[c]["tier_used"] == c[1]]
        scale = cm.calibrate([p for p, _ in fac], [m for _, m in fac])
        # a tier that fell back to FP64 in the subset costs what it cost there
        pred = {}
        for c in cands:
            fell = [m for (b, t), m in measured.items() if t == c[1] and info[(b, t)]["tier_used"] != t]
            pred[c] = float(np.median(fell)) if fell else scale * raw[c] + refine.get(c[1], 0.0)
        # measured candidates rank by their measurement, the rest by prediction
        pick = min(cands, key=lambda c: measured.get(c, pred[c]))
        if pick not in measured:
            t0 = time.perf_counter()
            measured[pick], _, info[pick] = sb.run(pick[0], pick[1])
            t_meas += time.perf_counter() - t0
        fallback = model_error(pred, measured) > max_error or abs(pred[pick] / measured[pick] - 1.0) > max_error
        if fallback:
            for c in cands:
                if c not in measured:
                    t0 = time.perf_counter()
                    measured[c], _, info[c] = sb.run(c[0], c[1])
                    t_meas += time.perf_counter() - t0
            pick = min(cands, key=measured.get)
        exhaustive, t_all = None, None
        if measure_all:
            exhaustive, t_all = {}, 0.0
            for c in cands:
                t0 = time.perf_counter()
                exhaustive[c] = sb.run(c[0], c[1])[0]
                t_all += time.perf_counter() - t0
        out = _report(cands, pred, measured, t_meas, pick, fallback, t_all, exhaustive)
        out.update({"algo": algo, "N": N, "train_samples": nsamp, "train_loss": [float(loss[0]), float(loss[-1])],
                    "scale": scale, "second": "tier", "refine_ms": refine})
        if knobs:
            (q, pr), km = tune_knobs(P, sb, pick[0], pick[1], N)
            out["knobs"] = {"sweep_block": q, "pair": pr,
                            "measured_ms": {f"q{kq}_pair{kp}": round(v, 3) for (kq, kp), v in km.items()}}
        return out
    finally:
        sb.free()


def tune_batch(P, n, kind, precision=0, grids=None, reps=3, seed=0):
    """Workgroups of the persistent batched launch (palial_opts.batch_grid):
    each candidate measured (device time of the launch, median of reps);
    candidates: 0 (as many as the CUs hold), and 1, 2 and 4 workgroups per CU."""
    grids = grids or (0, cm.NCU, 2 * cm.NCU, 4 * cm.NCU)
    bt = P.Batch(n, kind, precision=precision)
    try:
        res = {}
        for g in grids:
            ts = []
            for _ in range(reps + 1):
                bt.generate(seed)
                ts.append(bt.factor(solve=False, grid=g).factor_ms)
            res[g] = float(np.median(ts[1:]))
            if bt.info.any():
                raise RuntimeError(f"batched factorization failed at grid {g}")
        best = min(res, key=res.get)
        return {"grid": best, "measured_ms": {str(g): round(v, 3) for g, v in res.items()}}
    finally:
        bt.free()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--algo", default="cholesky", choices=("cholesky", "lu", "qr"))
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--solve", action="store_true", help="tune the solve (tile x tier) instead of the factorization")
    ap.add_argument("--kappa", type=float, default=1e2)
    ap.add_argument("--measure-all", action="store_true", help="also measure every candidate")
    args = ap.parse_args(argv)
    import palial_amd as P
    P.init(0)
    if args.solve:
        gen = P.CondGenerator(args.n, 512, seed=0)
        A = P.DeviceMatrix(args.n, 512)
        gen.fill(A, args.kappa)
        gen.free()
        try:
            res = tune_solve(P, A, args.algo, measure_all=args.measure_all)
        finally:
            A.free()
    else:
        res = tune_factor(P, args.algo, args.n, measure_all=args.measure_all)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
