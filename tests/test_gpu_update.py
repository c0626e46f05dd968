"""GPU tests of one trailing update (the O(n^3) part of a factorization step)
through the palial_hip_update test hook, in every precision tier.

The model is numpy in FP64 on operands rounded exactly as the device rounds
them (FP64 -> FP32 nearest, then BF16 nearest-even, or E4M3 nearest-even
after a per-tile power-of-two scale): the only remaining
difference is the FP32 accumulation order, bounded by b * 2^-24 * sum |a||b|.
The FP64 tier is checked with the FP64 unit roundoff in place of 2^-24.  The Cholesky update
touches only tiles i <= j (and the upper triangle of diagonal tiles), as
src/cholesky.h:86-170 does; the LU update touches every trailing tile
(src/lu.h:122-170)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def round_fp32(x):
    return x.astype(np.float32).astype(np.float64)


def round_bf16(x):
    u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split_bf16(x):
    """Split BF16 panels (precision 4): hi = BF16(FP32(x)), lo = BF16(FP32(x) - hi)."""
    hi = round_bf16(x)
    lo = round_bf16(round_fp32(x) - hi)
    return hi, lo


def round_e4m3(y):
    """E4M3 nearest-even of values below 128 in magnitude: 3 fraction bits,
    spacing 2^-9 below the smallest normal 2^-6."""
    a = np.abs(y)
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    q = 2.0 ** (np.maximum(e, -6.0) - 3.0)
    return np.where(a > 0, np.round(y / q) * q, y)


def round_fp8_tiles(P, b, axis):
    """FP8 panels (precision 3): each b x b tile of P (tiles along `axis`) is
    rounded to FP32, scaled by 2^e, e = 7 - frexp(max |tile|)[1], rounded to
    E4M3 and scaled back (both scalings exact)."""
    out = np.empty_like(P)
    for t in range(P.shape[axis] // b):
        sl = [slice(None), slice(None)]
        sl[axis] = slice(t * b, (t + 1) * b)
        f = round_fp32(P[tuple(sl)])
        mx = float(np.abs(f).max())
        e = 7 - np.frexp(mx)[1] if mx > 0 else 0
        out[tuple(sl)] = round_e4m3(f * 2.0 ** e) * 2.0 ** -e
    return out


ROUND = {0: lambda x: x, 1: round_fp32, 2: round_bf16}
TOL = {0: 4.0 * 2.0 ** -53, 1: 4.0 * 2.0 ** -24, 2: 4.0 * 2.0 ** -24, 3: 4.0 * 2.0 ** -24, 4: 4.0 * 2.0 ** -24}


def model_update(A, s0, s1, algo, prec):
    """Step-k update with operands as the device rounds them; the split tier
    sums hi*hi + hi*lo + lo*hi (lo*lo is not formed); the FP8 tier scales
    each panel tile by its own power of two."""
    b = s1 - s0
    if prec == 3:
        Ur = round_fp8_tiles(A[s0:s1, s1:], b, axis=1)
        Lr = Ur.T if algo == "cholesky" else round_fp8_tiles(A[s1:, s0:s1], b, axis=0)
        return Lr @ Ur, np.abs(Lr) @ np.abs(Ur)
    if algo == "cholesky":
        P = A[s0:s1, s1:]
        if prec == 4:
            h, l = split_bf16(P)
            return h.T @ h + h.T @ l + l.T @ h, np.abs(P).T @ np.abs(P)
        U = ROUND[prec](P)
        return U.T @ U, np.abs(U).T @ np.abs(U)
    Lc, Ur = A[s1:, s0:s1], A[s0:s1, s1:]
    if prec == 4:
        lh, ll = split_bf16(Lc)
        uh, ul = split_bf16(Ur)
        return lh @ uh + lh @ ul + ll @ uh, np.abs(Lc) @ np.abs(Ur)
    Lr, Urr = ROUND[prec](Lc), ROUND[prec](Ur)
    return Lr @ Urr, np.abs(Lr) @ np.abs(Urr)


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
@pytest.mark.parametrize("prec", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("N,b,k", [(1024, 256, 0), (1024, 256, 1), (1536, 512, 0), (768, 128, 2)])
def test_trailing_update(hip, orc, algo, prec, N, b, k):
    a = orc.gen("dense", N, b, seed=11 + k, shift=0.0)
    A = orc.to_dense(a, N, b)
    dm = hip.DeviceMatrix(N, b)
    try:
        dm.from_host(a)
        hip.update(algo, dm, k, precision=prec)
        out = orc.to_dense(dm.to_host(), N, b)
    finally:
        dm.free()
    s0, s1 = k * b, (k + 1) * b
    exp = A.copy()
    upd, scale = model_update(A, s0, s1, algo, prec)
    if algo == "cholesky":
        # upper triangle of the trailing block, tile by tile (i <= j)
        mask = np.zeros((N - s1, N - s1), dtype=bool)
        idx = np.arange(N - s1)
        ti = idx // b
        same = ti[:, None] == ti[None, :]
        mask |= ti[:, None] < ti[None, :]
        mask |= same & (idx[:, None] <= idx[None, :])
    else:
        mask = np.ones((N - s1, N - s1), dtype=bool)
    exp[s1:, s1:] = np.where(mask, A[s1:, s1:] - upd, A[s1:, s1:])
    err = np.abs(out - exp)
    # untouched regions are bitwise unchanged
    assert np.array_equal(out[:s1, :], A[:s1, :])
    assert np.array_equal(out[:, :s1], A[:, :s1])
    assert np.array_equal(out[s1:, s1:][~mask], A[s1:, s1:][~mask])
    bound = TOL[prec] * b * scale[mask] + 1e-300
    assert (err[s1:, s1:][mask] <= bound).all(), float((err[s1:, s1:][mask] / bound).max())


@pytest.mark.parametrize("algo", ["cholesky", "lu"])
@pytest.mark.parametrize("prec", [1, 2, 3, 4])
@pytest.mark.parametrize("N,b,k", [(1024, 256, 0), (1536, 512, 1)])
def test_trailing_update_fp32_storage(hip, orc, algo, prec, N, b, k):
    """The same update on FP32 tiles (the mixed-precision solves' working copy,
    hook flag prec | 8): operands rounded from the FP32 values, the result
    rounded to FP32 once more; untouched elements keep their FP32 bits."""
    a32 = orc.gen("dense", N, b, seed=21 + k, shift=0.0).astype(np.float32)
    buf = np.zeros(N * N, dtype=np.float64)
    buf.view(np.float32)[:N * N] = a32
    dm = hip.DeviceMatrix(N, b)
    try:
        dm.from_host(buf)
        hip.update(algo, dm, k, precision=prec | 8)
        out32 = dm.to_host().view(np.float32)[:N * N].copy()
    finally:
        dm.free()
    A = orc.to_dense(a32.astype(np.float64), N, b)
    out = orc.to_dense(out32.astype(np.float64), N, b)
    s0, s1 = k * b, (k + 1) * b
    upd, scale = model_update(A, s0, s1, algo, prec)
    if algo == "cholesky":
        idx = np.arange(N - s1)
        ti = idx // b
        mask = (ti[:, None] < ti[None, :]) | ((ti[:, None] == ti[None, :]) & (idx[:, None] <= idx[None, :]))
    else:
        mask = np.ones((N - s1, N - s1), dtype=bool)
    exp = A[s1:, s1:] - upd
    assert np.array_equal(out[:s1, :], A[:s1, :]) and np.array_equal(out[:, :s1], A[:, :s1])
    assert np.array_equal(out[s1:, s1:][~mask], A[s1:, s1:][~mask])
    err = np.abs(out[s1:, s1:] - exp)[mask]
    bound = TOL[prec] * b * scale[mask] + 2.0 ** -24 * np.abs(exp[mask]) + 1e-300
    assert (err <= bound).all(), float((err / bound).max())
