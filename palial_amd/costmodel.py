"""Task-time cost model: a 3-layer perceptron (64 hidden units, FP32) whose
forward pass and training run in the library's own HIP kernels
(palial_hip_mlp_forward / palial_hip_mlp_train), trained on per-launch device
times recorded by palial_hip_profile(2).

A launch is described by its kernel class, tile size, precision tier and
algorithmic flops and bytes; the model predicts log(ms).  `launches()`
enumerates the launches a factorization issues (the same batched launches
the library records), split into the diagonal chain and the bulk stream, so
a schedule's time can be predicted as the sum over steps of the slower of
the two."""
import ctypes as C
import math

import numpy as np

NCLASS = 7                 # kernel classes of palial_hip_profile_stats
NFEAT = NCLASS + 3 + 5     # class one-hot, log2 b, log10 flops, log10 bytes, tier one-hot
HID = 64
CLASSES = {"gemm": 0, "diag": 1, "trsm": 2, "tsqrt": 3, "qrupd": 4, "pack": 5, "pkgemm": 6}


def featurize(records):
    """records: (n, 6) array of (class, b, tier, flops, bytes, ms) -> (X float32 (n, NFEAT), y = log ms)."""
    r = np.asarray(records, dtype=np.float64).reshape(-1, 6)
    X = np.zeros((len(r), NFEAT), dtype=np.float32)
    cls = r[:, 0].astype(int)
    X[np.arange(len(r)), cls] = 1.0
    X[:, NCLASS] = np.log2(np.maximum(r[:, 1], 1.0)) / 9.0
    X[:, NCLASS + 1] = np.log10(r[:, 3] + 1.0) / 13.0
    X[:, NCLASS + 2] = np.log10(r[:, 4] + 1.0) / 11.0
    # tiers: precision 0 FP64, 1 FP32, 2 BF16, 3 FP8, 4 split BF16 -> columns 0..4
    tier = r[:, 2].astype(int)
    X[np.arange(len(r)), NCLASS + 3 + tier.clip(0, 4)] = 1.0
    y = np.log(np.maximum(r[:, 5], 1e-4)).astype(np.float32)
    return X, y


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


def launches(algo, N, b, prec=0, split=True, fs=None):
    """Launches of one factorization in the diagonal-chain schedule:
    list of (k, stream, class, flops, bytes) with stream 'chain', 'aux'
    (row/column TRSMs and packing of step k), 'first' (update of tile row
    k+1 and tile (k+2,k+2)) or 'rest' (the other trailing tiles).  Mirrors
    dag_cholesky_chain / dag_lu_chain (palial_amd/csrc/api.hip) with the split
    update (split=False: one update launch per step, labelled 'first'); the
    mixed tiers use packed panels when b % 128 == 0.  fs: the solve's FP32
    working copy (default: packed panels on the split schedule), with FP32
    tiles in the byte counts and the chain's diagonal update on the packed
    kernel."""
    MT = N // b
    b3, tb = float(b) ** 3, 8.0 * b * b
    packed = prec > 0 and b % 128 == 0
    if fs is None:
        fs = packed and split and MT > 1
    sbt = 4.0 * b * b if fs else tb             # bytes of one stored tile
    esz = {2: 2.0, 3: 1.0}.get(prec, 4.0)     # packed bytes per element (split BF16: hi + lo)
    out = []
    chol = algo == "cholesky"
    npan = 1 if chol else 2                      # packed panels per step
    for k in range(MT):
        T = MT - k - 1
        if chol:
            out.append((k, "chain", CLASSES["diag"], b3 / 3, 2 * tb))
        else:
            out.append((k, "chain", CLASSES["diag"], 2 * b3 / 3, 2 * tb))
        if T > 0:
            for _ in range(npan):
                out.append((k, "chain", CLASSES["trsm"], b3, 3 * sbt))
            if fs:
                for _ in range(npan):
                    out.append((k, "chain", CLASSES["pack"], 0.0, sbt + esz * b * b))
                out.append((k, "chain", CLASSES["pkgemm"], b3 if chol else 2 * b3, 2 * sbt + npan * esz * b * b))
            else:
                out.append((k, "chain", CLASSES["gemm"], b3 if chol else 2 * b3, 3 * tb))
        if T <= 1:
            continue
        for _ in range(npan):
            out.append((k, "aux", CLASSES["trsm"], (T - 1) * b3, 3.0 * (T - 1) * sbt))
        if packed:
            for _ in range(npan):
                out.append((k, "aux", CLASSES["pack"], 0.0, T * (sbt + esz * b * b)))
        # (gemm tiles, syrk tiles, packed panel tiles read) of each update launch
        if chol:
            parts = [("first", T - 1.0, 1.0, T), ("rest", (T - 1) * (T - 2) / 2.0, T - 2.0, T - 1)] if split \
                else [("first", T * (T - 1) / 2.0, T - 1.0, T)]
        else:
            parts = [("first", 2.0 * T - 1, 0.0, T), ("rest", (T - 1.0) ** 2 - 1, 0.0, T - 1)] if split \
                else [("first", T * T - 1.0, 0.0, T)]
        for st, ng, ns, np_ in parts:
            if ng + ns <= 0:
                continue
            fl = 2 * b3 * ng + b3 * ns
            if packed:
                out.append((k, st, CLASSES["pkgemm"], fl, 2.0 * (ng + ns) * sbt + npan * np_ * esz * b * b))
            else:
                out.append((k, st, CLASSES["gemm"], fl, 3.0 * (ng + ns) * tb))
    return out


def predict_factor_ms(model, algo, N, b, prec=0):
    """Predicted factorization time of the split chain schedule: step k's
    first update launch, then its remaining update beside step k+1's row
    TRSMs and packing, against the chain's step (both bounded below by it)."""
    L = launches(algo, N, b, prec)
    recs = np.array([(c, b, prec, f, by, 1.0) for (_, _, c, f, by) in L])
    X, _ = featurize(recs)
    ms = np.exp(model.predict_log_ms(X).astype(np.float64))
    MT = N // b
    acc = {st: np.zeros(MT + 1) for st in ("chain", "aux", "first", "rest")}
    for (k, st, _, _, _), t in zip(L, ms):
        acc[st][k] += t
    total = acc["aux"][0]
    for k in range(MT):
        bulk = acc["first"][k] + max(acc["rest"][k], acc["aux"][k + 1])
        total += max(acc["chain"][k], bulk)
    return float(total)
