"""GPU test of the layout-L dump: a device matrix written by the generator
kernel, dumped and loaded into a new device matrix, is the same bit for bit."""
import numpy as np
import pytest

from palial_amd import tileio

pytestmark = pytest.mark.gpu


def test_device_round_trip(hip, tmp_path):
    N, b = 512, 128
    A = hip.DeviceMatrix(N, b)
    B = None
    try:
        A.generate(hip.GEN_PALIAL, seed=7)
        p = str(tmp_path / "a.tiles")
        tileio.save_device(p, A)
        B = tileio.load_device(p)
        assert (B.N, B.b) == (N, b)
        assert np.array_equal(B.to_host().view(np.uint64), A.to_host().view(np.uint64))
    finally:
        A.free()
        if B is not None:
            B.free()
