"""Time and energy per CU policy: the GPU counterpart of the reference's DVFS
experiment (palial_power/server.c:69-84 caps each core's clock by the type of
the task it runs, gemm/potrf high and trsm/syrk low; exp/data holds its
per-task times under two governors).

A process may not change the GPU's clocks or power cap, but it may choose
the CUs each task class runs on: the mixed tiers' trailing updates (bulk),
row TRSMs and packing (aux) and diagonal chain (chain) have streams of their
own, and PALIAL_CU_POLICY gives each of them a CU mask
(palial_hip_cu_policy).  This module runs the same solve or factorization
under several policies, each in a fresh process (the masks are set when the
library starts), and samples the GPU's power sensor (amdgpu hwmon,
power1_average or power1_input in microwatts) meanwhile:

    python -m palial_amd.energy --algo cholesky --precision 2 --n 16384 --runs 5

One JSON line per policy: median milliseconds of the timed runs, mean power
and energy over the timed region (null without a readable sensor).
"""
import argparse
import glob
import json
import os
import statistics
import subprocess
import sys
import threading
import time

# name -> environment; the first entry is the library's default placement
POLICIES = {
    "default": {},                                    # bulk and aux off CUs 0-7 (PALIAL_RESERVE_CUS=8)
    "all_cus": {"PALIAL_RESERVE_CUS": "0"},           # no masks
    "chain_reserved": {"PALIAL_CHAIN_CUS": "1"},      # the chain on CUs 0-7 only
    "trsm_low": {"PALIAL_CU_POLICY": "bulk=8-255,aux=8-71,chain=0-7"},   # row TRSMs on a quarter
    "update_low": {"PALIAL_CU_POLICY": "bulk=8-191,aux=8-255"},          # trailing updates on 3/4
}


def find_power_sensor(root="/sys/class/drm"):
    """Path of the first amdgpu hwmon power file (microwatts), or None."""
    for mon in sorted(glob.glob(os.path.join(root, "card*", "device", "hwmon", "hwmon*"))):
        try:
            with open(os.path.join(mon, "name")) as f:
                if f.read().strip() != "amdgpu":
                    continue
        except OSError:
            continue
        for leaf in ("power1_average", "power1_input"):
            p = os.path.join(mon, leaf)
            try:
                with open(p) as f:
                    int(f.read().strip())
                return p
            except (OSError, ValueError):
                continue
    return None


def integrate(samples, t0, t1):
    """Energy (J) and mean power (W) of (t, watts) samples over [t0, t1]:
    trapezoids, with the power at t0 and t1 interpolated."""
    pts = sorted(samples)
    if len(pts) < 2 or t1 <= t0 or pts[0][0] > t0 or pts[-1][0] < t1:
        return None, None

    def at(t):
        for (ta, pa), (tb, pb) in zip(pts, pts[1:]):
            if ta <= t <= tb:
                return pa if tb == ta else pa + (pb - pa) * (t - ta) / (tb - ta)
        return None
    inner = [(t, p) for t, p in pts if t0 < t < t1]
    seq = [(t0, at(t0))] + inner + [(t1, at(t1))]
    e = sum((tb - ta) * (pa + pb) / 2 for (ta, pa), (tb, pb) in zip(seq, seq[1:]))
    return e, e / (t1 - t0)


class PowerSampler:
    """Reads a power file every `period` seconds on a thread: (monotonic s, W)."""

    def __init__(self, path, period=0.005):
        self.path, self.period = path, period
        self.samples = []
        self._stop = threading.Event()
        self._th = None

    def _run(self):
        while not self._stop.is_set():
            try:
                with open(self.path) as f:
                    w = int(f.read().strip()) * 1e-6
                self.samples.append((time.monotonic(), w))
            except (OSError, ValueError):
                pass
            self._stop.wait(self.period)

    def __enter__(self):
        if self.path:
            self._th = threading.Thread(target=self._run, daemon=True)
            self._th.start()
        return self

    def __exit__(self, *exc):
        self._stop.set()
        if self._th:
            self._th.join()


def child(a):
    """One process, one policy: warm-up, then the timed runs between two
    monotonic stamps printed as one JSON line."""
    import numpy as np
    from . import (CondGenerator, DeviceMatrix, GEN_DENSE, GEN_SYM, _check, factor_device, finalize, init, lib,
                   solve)
    import ctypes as C
    init(0)
    N, b = a.n, a.tile
    A = DeviceMatrix(N, b)
    times = []
    try:
        if a.factor:
            kind = GEN_SYM if a.algo == "cholesky" else GEN_DENSE

            def once():
                A.generate(kind, seed=0, shift=N)
                _, rep = factor_device(a.algo, A, no_graph=True)
                return rep.factor_ms
        else:
            gen = CondGenerator(N, b, seed=0)
            gen.fill(A, a.kappa)
            gen.free()
            rhs = np.random.default_rng(1).standard_normal(N)
            dX = C.c_void_p()
            _check(lib().palial_hip_malloc(C.byref(dX), N * 8))

            def once():
                _check(lib().palial_hip_memcpy_h2d(dX, rhs.ctypes.data, N * 8))
                _, rep, rc = solve(A, (dX, 1, N), algo=a.algo, precision=a.precision, device=True,
                                   factor_tile=b, check=False)
                if rc != 0:
                    raise RuntimeError(f"solve returned {rc}")
                return rep.factor_ms + rep.solve_ms
        once()
        t0 = time.monotonic()
        for _ in range(a.runs):
            times.append(once())
        t1 = time.monotonic()
    finally:
        A.free()
        finalize()
    print(json.dumps({"t0": t0, "t1": t1, "ms": times}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--algo", choices=["cholesky", "lu"], default="cholesky")
    ap.add_argument("--precision", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--tile", type=int, default=512)
    ap.add_argument("--kappa", type=float, default=1e2)
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--factor", action="store_true", help="FP64 factorizations instead of solves")
    ap.add_argument("--policies", default=",".join(POLICIES),
                    help="comma-separated names from: " + ", ".join(POLICIES) + "; or NAME=policy, the "
                         "PALIAL_CU_POLICY text with ';' between its entries")
    ap.add_argument("--child", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args(argv)
    if a.child:
        return child(a)
    sensor = find_power_sensor()
    rc = 0
    for item in a.policies.split(","):
        name, _, text = item.partition("=")
        env = dict(POLICIES[name]) if not text else {"PALIAL_CU_POLICY": text.replace(";", ",")}
        cmd = [sys.executable, "-m", "palial_amd.energy", "--child", "--algo", a.algo, "--precision",
               str(a.precision), "--n", str(a.n), "--tile", str(a.tile), "--kappa", str(a.kappa), "--runs",
               str(a.runs)] + (["--factor"] if a.factor else [])
        with PowerSampler(sensor) as ps:
            p = subprocess.run(cmd, env={**os.environ, **env}, capture_output=True, text=True,
                               cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        if p.returncode != 0:
            print(json.dumps({"policy": name, "env": env, "error": p.stderr.strip().splitlines()[-1:]}), flush=True)
            rc = 1
            continue
        r = json.loads(p.stdout.strip().splitlines()[-1])
        e, w = integrate(ps.samples, r["t0"], r["t1"])
        print(json.dumps({"policy": name, "env": env, "ms_median": round(statistics.median(r["ms"]), 3),
                          "runs": len(r["ms"]), "energy_j": None if e is None else round(e, 3),
                          "mean_w": None if w is None else round(w, 1), "sensor": sensor,
                          "workload": f"{a.algo} {'factorization' if a.factor else 'solve'} N={a.n} b={a.tile}"
                                      + ("" if a.factor else f" precision {a.precision} kappa {a.kappa:g}")}),
              flush=True)
    return rc


if __name__ == "__main__":
    sys.exit(main())
