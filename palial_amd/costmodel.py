"""Task-time cost model: a 3-layer perceptron (64 hidden units, FP32) whose
forward pass and training run in the library's own HIP kernels
(palial_hip_mlp_forward / palial_hip_mlp_train), trained on the per-task
device times that traced factorizations record (palial_hip_trace: the
reference's `name, ms` task times, measured on the GPU).

One sample is one reference tile task of one factorization.  Features
(NFEAT = 16): the task's kernel type (6: diagonal factorization, TRSM, SYRK,
GEMM, TSQRT, QR block-reflector update), log2 of the tile size, the precision
tier (5), log2 N, the task's DAG depth k/MT, log2 of the number of tasks in
its batched launch (its concurrency), and whether it lies on the critical
path (the diagonal chain).  The model predicts log(ms); a task's time is
taken to be the same under every schedule, which enters through the replay
below.

`predict_ms` turns per-task predictions into a factorization time by
replaying the library's schedule for that configuration: the streams of the
diagonal-chain schedule (or the QR panel chains and update streams) with
their cross-stream waits, each batched launch lasting as long as its tasks'
workgroups take on the CUs (sum over tasks of time x workgroups / resident
workgroup slots, and at least its longest task), the makespan bounded below
by the total work.  A per-algorithm scale is fitted on the measured
configurations (`calibrate`)."""
import ctypes as C
import math

import numpy as np

HID = 64
KTYPES = ("diag", "trsm", "syrk", "gemm", "tsqrt", "qrupd")
NTYPE = len(KTYPES)
NFEAT = NTYPE + 1 + 5 + 4          # 16: type one-hot, log2 b, tier one-hot, log2 N, depth, log2 concurrency, chain
# library task kinds (palial_hip_task_name) -> kernel type
KIND_TYPE = {0: 0, 1: 1, 2: 2, 3: 3, 4: 0, 5: 1, 6: 1, 7: 3, 8: 0, 9: 5, 10: 4, 11: 5}
ALGO_CODE = {"cholesky": 0, "lu": 1, "qr": 2}
NCU = 256                           # MI355X compute units
LAUNCH_MS = 0.005                   # gap between dependent launches on one stream


def launch_tasks(algo, MT, kind, k, m, n):
    """Tasks in the batched launch of the default schedule that covers each
    task (arrays of kind, k, m, n; dag_cholesky_chain / dag_lu_chain /
    dag_qr): 1 for the chain's one-tile launches, the rest of the panel for
    the row TRSM, the tile row (and column) k+1 plus tile (k+2,k+2) for the
    first update launch, every other trailing tile for the second, T = MT-k-1
    for the QR updates."""
    kind, k, m, n = (np.asarray(x, dtype=np.int64) for x in (kind, k, m, n))
    T = MT - k - 1
    out = np.ones(len(kind), dtype=np.int64)
    q = (kind == 9) | (kind == 11)
    out[q] = np.maximum(T[q], 1)
    tr = (kind == 1) | (kind == 5)
    out[tr] = np.where(m[tr] == k[tr] + 1, 1, np.maximum(T[tr] - 1, 1))
    tu = kind == 6
    out[tu] = np.where(n[tu] == k[tu] + 1, 1, np.maximum(T[tu] - 1, 1))
    up = (kind == 2) | (kind == 3) | (kind == 7)
    r, c = update_tile(algo, kind, m, n)
    first = (r == k + 1) | (c == k + 1) | ((r == k + 2) & (c == k + 2))
    chain = (r == k + 1) & (c == k + 1)
    if algo == 0:
        rest = np.maximum(T * (T + 1) // 2 - 1 - T, 1)
        fcount = T
    else:
        rest = np.maximum((T - 1) ** 2 - 1, 1)
        fcount = 2 * T - 1
    val = np.where(chain, 1, np.where(first, fcount, rest))
    out[up] = val[up]
    return out


def on_chain(algo, kind, k, m, n):
    """1 for the tasks on the DAG's critical path, the diagonal chain: the
    diagonal factorizations, the TRSMs of the tiles next to the diagonal and
    the update of tile (k+1,k+1); QR: GEQRT and the TSQRT panel chain."""
    kind, k, m, n = (np.asarray(x, dtype=np.int64) for x in (kind, k, m, n))
    out = np.isin(kind, (0, 4, 8, 10))
    idx = np.where(kind == 6, n, m)
    out |= np.isin(kind, (1, 5, 6)) & (idx == k + 1)
    r, c = update_tile(algo, kind, m, n)
    out |= np.isin(kind, (2, 3, 7)) & (r == k + 1) & (c == k + 1)
    return out.astype(np.float32)


def update_tile(algo, kind, m, n):
    """(row, column) tile an update task writes: Cholesky syrk(k,m) -> (m,m),
    gemm(k,m,n) -> (n,m); LU gemm(k,m,n) -> (m,n)."""
    kind, m, n = (np.asarray(x, dtype=np.int64) for x in (kind, m, n))
    if algo == 0:
        return np.where(kind == 2, m, n), m
    return m, n


def features(algo, N, b, tier, recs):
    """Feature rows of the reference tasks given as (kind, k, m, n) rows."""
    r = np.asarray(recs, dtype=np.float64).reshape(-1, 4)
    MT = N // b
    a = ALGO_CODE[algo]
    X = np.zeros((len(r), NFEAT), dtype=np.float32)
    kinds = r[:, 0].astype(int)
    types = np.array([KIND_TYPE[x] for x in range(12)])[kinds]
    X[np.arange(len(r)), types] = 1.0
    X[:, NTYPE] = math.log2(b) / 10.0
    X[:, NTYPE + 1 + int(tier)] = 1.0
    X[:, NTYPE + 6] = math.log2(N) / 16.0
    X[:, NTYPE + 7] = r[:, 1] / max(MT, 1)
    lt = launch_tasks(a, MT, kinds, r[:, 1].astype(int), r[:, 2].astype(int), r[:, 3].astype(int))
    X[:, NTYPE + 8] = np.log2(1.0 + lt) / 12.0
    X[:, NTYPE + 9] = on_chain(a, kinds, r[:, 1].astype(int), r[:, 2].astype(int), r[:, 3].astype(int))
    return X


def dataset(runs, max_per_run=4000, seed=0):
    """(X, y = log ms) from traced runs: (algo, N, b, tier, schedule, records)
    with records the library's (kind, k, m, n, ms, start) rows; at most
    max_per_run tasks of each run (sampled per kind), unmeasured tasks
    dropped."""
    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for algo, N, b, tier, sched, rec in runs:
        rec = np.asarray(rec, dtype=np.float64).reshape(-1, 6)
        rec = rec[rec[:, 4] > 0]
        if len(rec) > max_per_run:
            keep = []
            kinds = np.unique(rec[:, 0])
            per = max(1, max_per_run // len(kinds))
            for kd in kinds:
                idx = np.flatnonzero(rec[:, 0] == kd)
                keep.extend(rng.choice(idx, size=min(per, len(idx)), replace=False))
            rec = rec[np.sort(keep)]
        Xs.append(features(algo, N, b, tier, rec[:, :4]))
        ys.append(np.log(rec[:, 4]).astype(np.float32))
    X = np.concatenate(Xs)
    y = np.concatenate(ys)
    perm = rng.permutation(len(X))             # the trainer takes mini-batches in order
    return X[perm], y[perm]


def nparams(nfeat=NFEAT):
    return nfeat * HID + HID + HID * HID + HID + HID + 1


def init_params(nfeat=NFEAT, seed=0):
    """He-initialised weights, zero biases, in the library's parameter layout."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((nfeat, HID)) * math.sqrt(2.0 / nfeat)
    W2 = rng.standard_normal((HID, HID)) * math.sqrt(2.0 / HID)
    w3 = rng.standard_normal(HID) * math.sqrt(1.0 / HID)
    return np.concatenate([W1.ravel(), np.zeros(HID), W2.ravel(), np.zeros(HID), w3, [0.0]]).astype(np.float32)


def unpack(p, nfeat=NFEAT):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + nfeat * HID].reshape(nfeat, HID); o += nfeat * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    return W1, b1, W2, b2, w3, p[o]


def forward_ref(p, X):
    """FP64 reference forward pass."""
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ w3 + b3


def grad_ref(p, X, y):
    """FP64 reference gradient of the mean squared error, in the parameter layout."""
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    h1p = X @ W1 + b1
    h1 = np.maximum(h1p, 0.0)
    h2p = h1 @ W2 + b2
    h2 = np.maximum(h2p, 0.0)
    out = h2 @ w3 + b3
    d = 2.0 * (out - y) / len(y)
    gw3 = h2.T @ d
    gb3 = d.sum()
    dh2 = np.outer(d, w3) * (h2p > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (h1p > 0)
    gW1 = X.T @ dh1
    gb1 = dh1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


class MLP:
    """The cost model; forward and training run on the GPU."""

    def __init__(self, params=None, seed=0):
        self.params = np.ascontiguousarray(init_params(NFEAT, seed) if params is None else params, dtype=np.float32)

    def predict_log_ms(self, X):
        from . import lib, _check
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.empty(len(X), dtype=np.float32)
        _check(lib().palial_hip_mlp_forward(len(X), X.shape[1], X.ctypes.data_as(C.c_void_p),
                                            self.params.ctypes.data_as(C.c_void_p),
                                            out.ctypes.data_as(C.c_void_p), None))
        return out

    def gradient(self, X, y):
        from . import lib, _check
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty_like(self.params)
        _check(lib().palial_hip_mlp_train(len(X), X.shape[1], X.ctypes.data_as(C.c_void_p),
                                          y.ctypes.data_as(C.c_void_p), self.params.ctypes.data_as(C.c_void_p),
                                          0, C.c_float(0.0), None, g.ctypes.data_as(C.c_void_p), None))
        return g

    def train(self, X, y, epochs=200, lr=3e-3):
        """Adam on the mean squared error; returns the per-epoch loss."""
        from . import lib, _check
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = np.empty(epochs, dtype=np.float32)
        _check(lib().palial_hip_mlp_train(len(X), X.shape[1], X.ctypes.data_as(C.c_void_p),
                                          y.ctypes.data_as(C.c_void_p), self.params.ctypes.data_as(C.c_void_p),
                                          epochs, C.c_float(lr), loss.ctypes.data_as(C.c_void_p), None, None))
        return loss




# ------------------------------------------------------------------ schedule replay
def _slots(ktype, b, tier):
    """Resident workgroups of a task's kernel on the whole GPU and workgroups
    per task: 128x128 output blocks (two workgroups per CU) for the updates,
    one strip workgroup per 32 columns for the TRSMs (one per CU in FP64, two
    in FP32), one workgroup for a diagonal or panel factorization."""
    if ktype in ("gemm", "syrk", "qrupd"):
        blk = (b // 128) ** 2 if b % 128 == 0 else (b // 32) ** 2
        per_cu = 2 if b % 128 == 0 else 4
        if ktype == "syrk":
            blk = max(1, (int(math.sqrt(blk)) * (int(math.sqrt(blk)) + 1)) // 2)
        return NCU * per_cu, blk
    if ktype == "trsm":
        return NCU * (2 if tier > 0 else 1), max(1, b // 32)
    return NCU, 1


def _launch_ms(pred, ktype, b, tier):
    """Duration of a batched launch from its tasks' predicted times."""
    if len(pred) == 0:
        return 0.0
    slots, w = _slots(ktype, b, tier)
    return max(float(np.max(pred)), float(np.sum(pred)) * w / slots)


def task_keys(algo, MT):
    """(kind, k, m, n) of every reference task in creation order, as an int array."""
    from . import trace
    code = {"cholesky": 3, "lu": 7, "qr": 11}[algo]
    out = []
    for _, task in trace.tasks(algo, MT):
        c, k, m, n = trace._key(task)
        out.append((code if c is None else c, k, m, n))
    return np.array(out, dtype=np.int64).reshape(-1, 4)


def _group(gid, ngroups, ms):
    s = np.bincount(gid, weights=ms, minlength=ngroups)
    mx = np.zeros(ngroups)
    np.maximum.at(mx, gid, ms)
    cnt = np.bincount(gid, minlength=ngroups)
    return s, mx, cnt


def replay(algo, N, b, tier, schedule, ms, keys=None):
    """Makespan (ms) of one factorization from per-task times ms (creation
    order), replaying the library's streams for the schedule: Cholesky / LU
    1 diagonal chain + split update (also 0, the default), 2 diagonal chain +
    one update launch, 3 row look-ahead (the panel stream updates tile row
    k+1, factors tile (k+1,k+1) and solves panel k+1 while the main stream
    updates the rest), 4 one stream; QR:
    schedule = panel chains in flight (0: 8)."""
    MT = N // b
    keys = task_keys(algo, MT) if keys is None else keys
    ms = np.asarray(ms, dtype=np.float64)
    kind, k, m, n = keys.T
    if algo == "qr":
        sl, w = _slots("qrupd", b, tier)
        ns = schedule if schedule > 0 else 8
        geq = np.zeros(MT)
        geq[k[kind == 8]] = ms[kind == 8] + LAUNCH_MS
        tsq = np.zeros((MT, MT))
        sel = kind == 10
        tsq[k[sel], m[sel]] = ms[sel] + LAUNCH_MS
        sel = kind == 9
        o_s, o_m, _ = _group(k[sel], MT, ms[sel])
        sel = kind == 11
        g_s, g_m, _ = _group(k[sel] * MT + m[sel], MT * MT, ms[sel])
        dur = lambda s_, m_: 3 * LAUNCH_MS + max(m_, s_ * w / sl)     # three launches per update
        chain_free = [0.0] * ns
        upd_free = [0.0] * ns
        tsm_prev = np.zeros(MT)
        for kk in range(MT):
            c = kk % ns
            g_end = max(chain_free[c], tsm_prev[kk] if kk > 0 else 0.0) + geq[kk]
            chain_free[c] = g_end
            tsm_now = np.zeros(MT)
            if kk + 1 < MT:
                upd_free[c] = max(upd_free[c], g_end) + dur(o_s[kk], o_m[kk])
                for mm in range(kk + 1, MT):
                    q_end = max(chain_free[c], tsm_prev[mm] if kk > 0 else 0.0) + tsq[kk, mm]
                    chain_free[c] = q_end
                    gi = kk * MT + mm
                    upd_free[c] = max(upd_free[c], q_end) + dur(g_s[gi], g_m[gi])
                    tsm_now[mm] = upd_free[c]
            tsm_prev = tsm_now
        work = (np.sum(ms[kind == 9]) + np.sum(ms[kind == 11])) * w / sl
        return max(max(chain_free), max(upd_free), work)
    chol = algo == "cholesky"
    dsel = kind == (0 if chol else 4)
    diag = np.zeros(MT)
    diag[k[dsel]] = ms[dsel] + LAUNCH_MS
    slt, wt = _slots("trsm", b, tier)
    slg, wg = _slots("gemm", b, tier)

    def launches(mask, slots, w):
        """Per-step duration of one batched launch holding the masked tasks
        (with its dispatch gap; 0 for a step without such tasks)."""
        s_, m_, cnt = _group(k[mask], MT, ms[mask])
        return np.where(cnt > 0, np.maximum(m_, s_ * w / slots) + LAUNCH_MS, 0.0)

    idx1 = np.where(kind == 6, n, m)
    # TRSM launches: Cholesky one panel; LU the U row (kind 6) and the L column (kind 5)
    tr_groups = [kind == 1] if chol else [kind == 6, kind == 5]
    tr_chain = sum(launches(g & (idx1 == k + 1), slt, wt) for g in tr_groups)
    tr_rest = sum(launches(g & (idx1 > k + 1), slt, wt) for g in tr_groups)
    tr_all = sum(launches(g, slt, wt) for g in tr_groups)
    up = (kind == 2) | (kind == 3) | (kind == 7)
    r, c = update_tile(0 if chol else 1, kind, m, n)
    s1 = up & (r == k + 1) & (c == k + 1)
    row = up & ((r == k + 1) | (c == k + 1))                 # tile row (LU: and column) k+1, with (k+1,k+1)
    first = up & ~s1 & (row | ((r == k + 2) & (c == k + 2)))
    rest = up & ~s1 & ~first
    cs = launches(s1, slg, wg)
    work = float(np.sum(ms[up]) * wg / slg)
    end = 0.0
    if schedule == 4:
        return float(np.sum(diag) + np.sum(tr_all) + np.sum(launches(up, slg, wg)))
    if schedule == 3:
        d_row = launches(row, slg, wg)
        d_rest = launches(up & ~row, slg, wg)
        s1_free = diag[0] + tr_all[0]
        p_ready = s1_free
        s0_free = 0.0
        u_prev = 0.0
        for kk in range(MT - 1):
            s1_free = max(s1_free, u_prev) + d_row[kk] + diag[kk + 1] + tr_all[kk + 1]
            s0_free = max(s0_free, p_ready) + d_rest[kk]
            u_prev = s0_free
            p_ready = s1_free
            end = max(end, s1_free, s0_free)
        return max(end, work)
    split = schedule in (0, 1)
    dfi = launches(first, slg, wg) if split else launches(first | rest, slg, wg)
    dre = launches(rest, slg, wg) if split else np.zeros(MT)
    p_end = diag[0]
    f_prev = r_prev = 0.0
    aux_free = bulk_free = 0.0
    end = p_end
    for kk in range(MT - 1):
        t1_end = max(p_end, f_prev if split else r_prev) + tr_chain[kk]
        p_next = t1_end + cs[kk] + diag[kk + 1]
        if split:
            aux_end = max(aux_free, f_prev, p_end) + tr_rest[kk]
            aux_free = aux_end
            f_end = max(bulk_free, aux_end, t1_end) + dfi[kk]
            bulk_free = f_end + dre[kk]
            f_prev, r_prev = f_end, bulk_free
        else:
            tb_end = max(bulk_free, p_end) + tr_rest[kk]
            bulk_free = max(tb_end, t1_end) + dfi[kk]
            f_prev = r_prev = bulk_free
        p_end = p_next
        end = max(end, p_end, bulk_free)
    return max(end, work)


def predict_ms(model, algo, N, b, tier, schedule):
    """Predicted factorization time (before calibration) of one configuration."""
    keys = task_keys(algo, N // b)
    X = features(algo, N, b, tier, keys)
    ms = np.exp(model.predict_log_ms(X).astype(np.float64))
    return replay(algo, N, b, tier, schedule, ms, keys)


def calibrate(pred, meas):
    """Scale c minimising the squared log error of c * pred against meas."""
    pred, meas = np.asarray(pred, dtype=np.float64), np.asarray(meas, dtype=np.float64)
    return float(np.exp(np.mean(np.log(meas) - np.log(pred))))
