"""CPU tests of the drop-in boundary: the C-ABI library loads, exports every
function include/palial_hip.h declares, keeps the reference descriptor layout,
and rejects misuse with error codes (no GPU work is issued here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "palial_hip.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(palial_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import palial_amd
    lib = palial_amd.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", palial_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (palial_hip_\w+)", out))
    assert set(names) <= exported


def test_descriptor_layout_matches_reference():
    # include/descriptor.h:13-20 on x86-64: offsets 0/8/16/24/32, sizeof 40
    import palial_amd
    import oracle
    for S in (palial_amd.MATRIX_desc, oracle.Desc):
        assert C.sizeof(S) == 40
        assert [getattr(S, f).offset for f, _ in S._fields_] == [0, 8, 16, 24, 32]


def test_calls_before_init_fail_cleanly():
    import palial_amd
    lib = palial_amd.lib()
    d = palial_amd.MATRIX_desc(64, 128 * 128, 64 * 64, 128, None)
    if palial_amd._inited:
        pytest.skip("library already initialised in this process")
    assert lib.palial_hip_cholesky(d, None, None) == -5          # PALIAL_ERR_STATE
    assert b"init" in lib.palial_hip_last_error()
    assert lib.palial_hip_strerror(-1) == b"invalid argument"


def test_driver_built():
    import palial_amd
    assert os.path.exists(palial_amd.DRIVER_PATH) or not os.path.exists(
        os.path.join(ROOT, "palial_amd", "driver", "palial.c"))


def test_matrix_arguments_are_checked():
    # the Python mirror refuses what the C ABI could not describe, before any call
    import numpy as np
    import palial_amd
    with pytest.raises(TypeError):
        palial_amd._desc_of(np.zeros(16, dtype=np.float32), b=2)
    with pytest.raises(TypeError):
        palial_amd._desc_of(np.zeros((8, 8))[:, ::2].copy(order="F")[:, :4], b=2)
    with pytest.raises(ValueError):
        palial_amd._desc_of(np.zeros(15), b=2)
    with pytest.raises(ValueError):
        palial_amd._desc_of(np.zeros(16))
    d, dev, _ = palial_amd._desc_of(np.zeros(16), b=2)
    assert (d.matrix_size, d.tile_size, d.matrix_nelements, d.tile_nelements, dev) == (4, 2, 16, 4, False)
