import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def gpu_available():
    try:
        from poe2d_amd import _abi
        lib = _abi.load_library()
        import ctypes
        n = ctypes.c_int(0)
        arch = ctypes.create_string_buffer(64)
        return lib.poe2d_device_info(ctypes.byref(n), arch, 64) == 0 and n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    return pyoracle.load()
