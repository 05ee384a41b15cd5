import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_present() -> bool:
    """True when a ROCm GPU device node exists (no HIP initialisation)."""
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def hip():
    """libnlhip.so with the HIP extension initialised (GPU tests only)."""
    from graphitethree_amd import capi
    L = capi.lib()
    if not L.nlInitExtension(b"HIP"):
        pytest.fail("nlInitExtension(\"HIP\") failed: " + capi.last_error())
    return L
