"""Dump and load of layout-L matrices, so fixtures, device results and the
CPU oracle can be compared offline.

File format (little-endian):
  bytes  0..7   magic b"PALIALT1"
  bytes  8..15  N (uint64)
  bytes 16..23  b (uint64), b divides N
  bytes 24..31  element size in bytes (uint64): 8 FP64, 4 FP32
  bytes 32..39  layout (uint64): 0 = layout L, the reference's tile order
                (include/tile_address.h:15: tile (m,n) at b*b*(m + (N/b)*n),
                column-major inside)
  bytes 40..63  zero
  bytes 64..    N*N elements in that order (the bytes a MATRIX_desc's
                `matrix` buffer holds)
"""
import os

import numpy as np

MAGIC = b"PALIALT1"
HEADER = 64
_DTYPES = {8: np.float64, 4: np.float32}


def save(path, a, N, b):
    """Write the N*N layout-L elements of `a` (FP64 or FP32) to `path`."""
    a = np.ascontiguousarray(a).reshape(-1)
    if a.dtype not in (np.float64, np.float32):
        raise TypeError(f"layout-L dump holds FP64 or FP32 elements, not {a.dtype}")
    if N <= 0 or b <= 0 or N % b:
        raise ValueError(f"tile size {b} does not divide N={N}")
    if a.size < N * N:
        raise ValueError(f"{a.size} elements for an {N} x {N} matrix")
    hdr = np.zeros(HEADER // 8, dtype="<u8")
    hdr[1:5] = (N, b, a.dtype.itemsize, 0)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(hdr[1:].tobytes())
        f.write(a[:N * N].astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


def read_header(path):
    """(N, b, dtype) of a dump; raises ValueError on a file that is not one."""
    with open(path, "rb") as f:
        raw = f.read(HEADER)
    if len(raw) < HEADER or raw[:8] != MAGIC:
        raise ValueError(f"{path}: not a layout-L dump")
    N, b, es, layout = (int(x) for x in np.frombuffer(raw[8:40], dtype="<u8"))
    if es not in _DTYPES or layout != 0 or b <= 0 or N <= 0 or N % b:
        raise ValueError(f"{path}: bad header (N={N}, b={b}, element size {es}, layout {layout})")
    if os.path.getsize(path) != HEADER + N * N * es:
        raise ValueError(f"{path}: {os.path.getsize(path)} bytes, expected {HEADER + N * N * es}")
    return N, b, np.dtype(_DTYPES[es]).newbyteorder("<")


def load(path, mmap=False):
    """(a, N, b): the flat layout-L elements (memory-mapped read-only with mmap=True)."""
    N, b, dt = read_header(path)
    if mmap:
        return np.memmap(path, dtype=dt, mode="r", offset=HEADER, shape=(N * N,)), N, b
    return np.fromfile(path, dtype=dt, offset=HEADER, count=N * N), N, b


def save_device(path, dm):
    """Dump a DeviceMatrix (FP64)."""
    save(path, dm.to_host(), dm.N, dm.b)


def load_device(path):
    """A new DeviceMatrix holding a FP64 dump."""
    from . import DeviceMatrix
    a, N, b = load(path)
    if a.dtype.itemsize != 8:
        raise ValueError(f"{path}: a DeviceMatrix holds FP64 elements")
    dm = DeviceMatrix(N, b)
    dm.from_host(np.ascontiguousarray(a, dtype=np.float64))
    return dm
