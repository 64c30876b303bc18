import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    try:
        import gpu_index_test_amd._native as N
        import ctypes
        c = ctypes.c_int(0)
        N.lib().gis_device_count(ctypes.byref(c))
        return c.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    import gpu_index_test_amd as G
    if not gpu_available():
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    return G.default_context()
