"""GPU test of palial_hip_dtsmqr_noalloc, the TSMQR tile kernel whose
workspace lives in a static device array (no allocation inside the call).

Its output must be bitwise that of palial_hip_dtsmqr (same kernels, same
inputs) on two calls in a row, the second with a smaller tile, and both must
match the oracle's PALIAL_dtsmqr (src/srcqr/dtsmqr.h:11) to 1e-12 relative.
Calls the C ABI directly through ctypes with library-allocated device tiles.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LEFT, TRANS, NOTRANS = 141, 112, 111


def tiles(n, seed):
    r = np.random.default_rng(seed)
    A1 = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    A2 = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    V = np.asfortranarray(r.uniform(-0.5, 0.5, (n, n)))
    T = np.asfortranarray(np.triu(r.uniform(-0.5, 0.5, (n, n))) + np.tril(r.uniform(-9, 9, (n, n)), -1))
    return A1, A2, V, T


class Dev:
    """Device copies of column-major tiles, allocated by the test (not by the
    entry point under test)."""

    def __init__(self, hip, arrays):
        self.hip, self.ptrs = hip, []
        lib = hip.lib()
        for a in arrays:
            p = C.c_void_p()
            assert lib.palial_hip_malloc(C.byref(p), a.nbytes) == 0
            assert lib.palial_hip_memcpy_h2d(p, a.ctypes.data, a.nbytes) == 0
            self.ptrs.append(p)

    def get(self, i, shape):
        out = np.empty(shape, order="F")
        assert self.hip.lib().palial_hip_memcpy_d2h(out.ctypes.data, self.ptrs[i], out.nbytes) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.lib().palial_hip_free(p)


def run(hip, fn, n, A1, A2, V, T, stream=None):
    d = Dev(hip, [A1, A2, V, T])
    o = hip._opts(device=True)
    o.stream = stream
    rc = fn(LEFT, TRANS, n, n, n, n, d.ptrs[0], n, d.ptrs[1], n, d.ptrs[2], n, d.ptrs[3], n, None, n, C.byref(o))
    if stream is not None:
        hip.synchronize()
    out = (d.get(0, (n, n)), d.get(1, (n, n)))
    d.free()
    return rc, out


def test_noalloc_matches_dtsmqr_twice(hip, orc):
    lib = hip.lib()
    for n, seed in ((512, 1), (256, 2)):     # second call: a smaller tile in the same scratch
        A1, A2, V, T = tiles(n, seed)
        print(f"dtsmqr_noalloc n={n} seed={seed}", flush=True)
        rc0, (r1, r2) = run(hip, lib.palial_hip_dtsmqr, n, A1, A2, V, T)
        rc1, (s1, s2) = run(hip, lib.palial_hip_dtsmqr_noalloc, n, A1, A2, V, T)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(r1.view(np.uint64), s1.view(np.uint64))
        assert np.array_equal(r2.view(np.uint64), s2.view(np.uint64))
        o1, o2 = orc.tile_tsmqr(A1, A2, V, T)
        for x, ref in ((s1, o1), (s2, o2)):
            assert np.abs(x - ref).max() / np.abs(ref).max() < 1e-12


def test_noalloc_on_caller_stream(hip, orc):
    """With opts->stream set the call only enqueues; after a device
    synchronize the result equals palial_hip_dtsmqr's."""
    if os.environ.get("PALIAL_HIPEMU"):
        pytest.skip("the CPU build of the library has no HIP runtime whose streams the test could create")
    lib = hip.lib()
    n = 128
    A1, A2, V, T = tiles(n, 3)
    _, (r1, r2) = run(hip, lib.palial_hip_dtsmqr, n, A1, A2, V, T)
    # a stream of the HIP runtime the library itself is linked against
    rt = C.CDLL(hip_runtime_path())
    st = C.c_void_p()
    assert rt.hipStreamCreate(C.byref(st)) == 0
    try:
        rc, (s1, s2) = run(hip, lib.palial_hip_dtsmqr_noalloc, n, A1, A2, V, T, stream=st)
    finally:
        rt.hipStreamDestroy(st)
    assert rc == 0
    assert np.array_equal(r1.view(np.uint64), s1.view(np.uint64))
    assert np.array_equal(r2.view(np.uint64), s2.view(np.uint64))


def hip_runtime_path():
    """Path of the libamdhip64 this process has mapped for libpalial_hip.so."""
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert len(paths) == 1, paths
    return paths.pop()


def test_noalloc_rejects_what_it_cannot_hold(hip):
    lib = hip.lib()
    dummy = C.c_void_p(0x1000)    # never dereferenced: the checks come first
    o = hip._opts(device=True)
    n = 1024
    rc = lib.palial_hip_dtsmqr_noalloc(LEFT, TRANS, n, n, n, n, dummy, n, dummy, n, dummy, n, dummy, n, None, n,
                                       C.byref(o))
    assert rc == -1 and b"512" in lib.palial_hip_last_error()
    host = np.zeros((64, 64), order="F")
    o.device_matrix = 0
    p = host.ctypes.data
    rc = lib.palial_hip_dtsmqr_noalloc(LEFT, TRANS, 64, 64, 64, 64, p, 64, p, 64, p, 64, p, 64, None, 64,
                                       C.byref(o))
    assert rc == -1 and b"device" in lib.palial_hip_last_error()
    o.device_matrix = 1
    rc = lib.palial_hip_dtsmqr_noalloc(LEFT, NOTRANS, 64, 64, 64, 64, dummy, 64, dummy, 64, dummy, 64, dummy, 64,
                                       None, 64, C.byref(o))
    assert rc == -1
