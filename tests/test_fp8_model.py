"""CPU check of the E4M3 rounding model that the FP8 update tests compare the
device against (test_gpu_update.round_e4m3).  The model is pinned to the
format's definition, not to itself: every finite E4M3 value is decoded from
its bit pattern (sign, 4 exponent bits with bias 7, 3 fraction bits,
subnormals at exponent field 0), and the model must keep each representable
value below 128, send each midpoint between neighbours to the one with the
even fraction, and send points just off a midpoint to the nearer neighbour."""
import numpy as np

from test_gpu_update import round_e4m3, round_fp8_tiles


def e4m3_values():
    """The non-negative finite E4M3 values, from the encoding (S.EEEE.MMM)."""
    vals = []
    for e in range(16):
        for m in range(8):
            if e == 15 and m == 7:
                continue                                   # NaN
            vals.append((m / 8.0) * 2.0 ** -6 if e == 0 else (1 + m / 8.0) * 2.0 ** (e - 7))
    return np.array(sorted(vals))


def test_decoded_grid():
    v = e4m3_values()
    assert v[0] == 0 and v[1] == 2.0 ** -9 and v[-1] == 448.0 and len(v) == 127
    assert len(np.unique(v)) == len(v)


def test_representable_values_are_fixed_points():
    v = e4m3_values()
    v = v[v < 128]
    assert np.array_equal(round_e4m3(v), v)
    assert np.array_equal(round_e4m3(-v), -v)


def test_midpoints_round_to_even_fraction():
    v = e4m3_values()
    v = v[v <= 128]
    lo, hi = v[:-1], v[1:]
    mid = (lo + hi) / 2
    # the even neighbour: fraction field even <=> value / spacing is even
    spacing = hi - lo
    even = np.where(np.round(lo / spacing) % 2 == 0, lo, hi)
    assert np.array_equal(round_e4m3(mid), even)
    eps = spacing * 2.0 ** -6
    assert np.array_equal(round_e4m3(mid - eps), lo)
    assert np.array_equal(round_e4m3(mid + eps), hi)


def test_tile_scaling_is_a_power_of_two_per_tile():
    """Each tile's largest magnitude lands in [64, 128) before rounding, so
    the rounding is relative (error <= 2^-4 of the value for normals)."""
    rng = np.random.default_rng(0)
    P = np.hstack([rng.standard_normal((8, 8)) * s for s in (1e-3, 1.0, 7e4)])
    Q = round_fp8_tiles(P, 8, axis=1)
    for t in range(3):
        p, q = P[:, 8 * t:8 * t + 8], Q[:, 8 * t:8 * t + 8]
        big = np.abs(p) >= np.abs(p).max() * 2.0 ** -6      # normal range after the scale
        assert (np.abs(q - p)[big] <= 2.0 ** -4 * np.abs(p)[big] * (1 + 1e-6)).all()
