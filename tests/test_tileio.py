"""CPU tests of the layout-L dump format (palial_amd.tileio): round trips
bit for bit in FP64 and FP32, the header, memory-mapped loads, the oracle's
generated tiles, and rejection of files that are not dumps."""
import numpy as np
import pytest

from palial_amd import tileio


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_round_trip_bitwise(tmp_path, dtype):
    N, b = 96, 32
    a = np.random.default_rng(0).standard_normal(N * N).astype(dtype)
    a[5] = -0.0
    a[7] = np.inf
    p = str(tmp_path / "m.tiles")
    tileio.save(p, a, N, b)
    assert (tmp_path / "m.tiles").stat().st_size == 64 + N * N * a.itemsize
    assert tileio.read_header(p)[:2] == (N, b)
    for mm in (False, True):
        got, n2, b2 = tileio.load(p, mmap=mm)
        assert (n2, b2) == (N, b) and got.dtype.itemsize == a.itemsize
        assert np.array_equal(np.asarray(got).view(np.uint8), a.view(np.uint8))


def test_oracle_tiles_round_trip(tmp_path, orc):
    N, b = 128, 32
    a = orc.gen("sym", N, b, seed=3, shift=float(N))
    p = str(tmp_path / "s.tiles")
    tileio.save(p, a, N, b)
    got, _, _ = tileio.load(p)
    assert np.array_equal(got, a)
    D = orc.to_dense(got, N, b)
    assert np.array_equal(D, D.T)                       # the symmetric generator, read back in layout L


def test_rejects_non_dumps(tmp_path):
    N, b = 64, 32
    p = str(tmp_path / "x.tiles")
    tileio.save(p, np.zeros(N * N), N, b)
    raw = bytearray(open(p, "rb").read())
    bad = tmp_path / "bad"
    bad.write_bytes(b"NOTATILE" + raw[8:])
    with pytest.raises(ValueError):
        tileio.load(str(bad))
    bad.write_bytes(raw[:-8])                           # truncated
    with pytest.raises(ValueError):
        tileio.load(str(bad))
    raw2 = bytearray(raw)
    raw2[16:24] = np.array([48], dtype="<u8").tobytes()  # b does not divide N
    bad.write_bytes(bytes(raw2))
    with pytest.raises(ValueError):
        tileio.load(str(bad))
    with pytest.raises(ValueError):
        tileio.save(p, np.zeros(10), N, b)
    with pytest.raises(TypeError):
        tileio.save(p, np.zeros(N * N, dtype=np.int64), N, b)
