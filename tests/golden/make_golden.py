"""Write the golden fixtures under tests/golden/ (run in the build container).

1. chol_task_names.json -- the Cholesky task order of the reference's own
   per-task traces (exp/data/tasks_times_<N>_512_*: one "name, ms" line per
   task in creation order, palial_trace/log.h:47-48).  Stored: the full name
   list for MT=16 and a sha256 of the name sequence for MT in {16,32,48,64}.
   Only the names (data) are kept; the timings are summarised in BASELINE.md.
2. lapack_tiles.npz -- tile-kernel input/output vectors from LAPACK via scipy
   (dpotrf, dgeqrt, dtpqrt(l=0), dtpmqrt(l=0,'L','T'), dgemqrt('L','T'),
   dtrsm/dsyrk/dgemm forms) at b in {4, 8, 32}; independent of the oracle.

    python tests/golden/make_golden.py
"""
import glob
import hashlib
import json
import os

import numpy as np
import scipy.linalg as sl
import scipy.linalg.lapack as la

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/exp/data"


def task_names():
    out = {"source": "exp/data/tasks_times_<N>_512_<T>_<governor> (reference traces)", "sha256": {},
           "names_MT16": None}
    for f in sorted(glob.glob(os.path.join(REF, "tasks_times_*"))):
        N = int(os.path.basename(f).split("_")[2])
        names = [l.split(",")[0].strip() for l in open(f) if l.strip()]
        h = hashlib.sha256("\n".join(names).encode()).hexdigest()
        MT = N // 512
        prev = out["sha256"].get(str(MT))
        assert prev in (None, h), f"traces for MT={MT} disagree on task order"
        out["sha256"][str(MT)] = h
        if MT == 16:
            out["names_MT16"] = names
    json.dump(out, open(os.path.join(HERE, "chol_task_names.json"), "w"))


def lapack_tiles():
    rng = np.random.default_rng(2024)
    d = {}
    for n in (4, 8, 32):
        a = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)))
        spd = np.asfortranarray((a + a.T) / 2 + n * np.eye(n))
        d[f"potrf_in_{n}"] = spd
        d[f"potrf_out_{n}"] = sl.cholesky(spd, lower=False)
        g = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)))
        qr, t, info = la.dgeqrt(n, g)
        d[f"geqrt_in_{n}"] = g; d[f"geqrt_A_{n}"] = qr; d[f"geqrt_T_{n}"] = np.triu(t)
        a2 = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)))
        r = np.asfortranarray(np.triu(qr))
        r1, b1, t1, info = la.dtpqrt(0, n, r.copy(order="F"), a2.copy(order="F"))
        d[f"tsqrt_R_{n}"] = r; d[f"tsqrt_A2_{n}"] = a2
        d[f"tsqrt_A1out_{n}"] = r1; d[f"tsqrt_A2out_{n}"] = b1; d[f"tsqrt_T_{n}"] = np.triu(t1)
        c1 = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)))
        c2 = np.asfortranarray(rng.uniform(-0.5, 0.5, (n, n)))
        o1, o2, info = la.dtpmqrt(0, b1, t1, c1.copy(order="F"), c2.copy(order="F"), side="L", trans="T")
        d[f"tsmqr_C1_{n}"] = c1; d[f"tsmqr_C2_{n}"] = c2; d[f"tsmqr_out1_{n}"] = o1; d[f"tsmqr_out2_{n}"] = o2
        cm, info = la.dgemqrt(qr, t, c1.copy(order="F"), side="L", trans="T")
        d[f"ormqr_C_{n}"] = c1; d[f"ormqr_out_{n}"] = cm
    np.savez_compressed(os.path.join(HERE, "lapack_tiles.npz"), **d)


if __name__ == "__main__":
    if os.path.isdir(REF):
        task_names()
    lapack_tiles()
    print("golden fixtures written to", HERE)
