import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available() -> bool:
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def ref_driver():
    """Path of the reference build (oracle/_ref/ref_driver), or skip."""
    p = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(p):
        pytest.skip("oracle/_ref/ref_driver not built")
    return p


@pytest.fixture(scope="session")
def shim_driver():
    """Path of the reference's set-up and driver loop with the solver entry
    points of minifs_amd/shim on libmfs_hip.so (oracle/_ref/shim_driver), or skip."""
    p = os.path.join(ROOT, "oracle", "_ref", "shim_driver")
    if not os.path.exists(p):
        pytest.skip("oracle/_ref/shim_driver not built")
    return p
