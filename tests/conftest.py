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
    # PALIAL_HIPEMU=plain|asan: the GPU tests run on the CPU build of the same
    # sources (tests/emu); the asan variant needs the ASan runtime preloaded
    variant = os.environ.get("PALIAL_HIPEMU")
    if variant:
        from emu import build as emu_build
        palial_amd.LIB_PATH = emu_build.build(variant)
    palial_amd.init(0)
    return palial_amd
