import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)     # test helpers (layout_probe)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hip():
    import palial_amd
    # PALIAL_HIPEMU=plain|asan|race: the GPU tests run on the CPU build of the
    # same sources (tests/emu); the asan variant needs the ASan runtime
    # preloaded, the race variant checks every device-memory access against
    # the order of the stream operations
    variant = os.environ.get("PALIAL_HIPEMU")
    if variant:
        from emu import build as emu_build
        palial_amd.LIB_PATH = emu_build.build(variant)
    palial_amd.init(0)
    return palial_amd


@pytest.fixture(autouse=True)
def _no_stream_races(request):
    """Race-checking CPU build (PALIAL_HIPEMU=race): a test fails if any of
    its device-memory accesses was unordered with a conflicting one."""
    yield
    if os.environ.get("PALIAL_HIPEMU") != "race" or "hip" not in request.fixturenames:
        return
    import ctypes as C
    import palial_amd
    L = palial_amd.lib()
    L.hipemu_races.restype = C.c_long
    L.hipemu_race_report.restype = C.c_long
    n = L.hipemu_races()
    if n:
        buf = C.create_string_buffer(1 << 16)
        L.hipemu_race_report(buf, len(buf))
        L.hipemu_race_reset()
        pytest.fail(f"{n} racing accesses:\n{buf.value.decode()}")
