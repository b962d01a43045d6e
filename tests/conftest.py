import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when a gpu-marked test runs without a GPU or without the
    built library: the HIP path is the only path."""
    import openatom_amd
    openatom_amd.load()
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    rc = hip.hipGetDeviceCount(ctypes.byref(n))
    if rc != 0 or n.value < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return n.value
