import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def gpu_available() -> bool:
    """True when the library can open a context on device 0 (torch is not
    imported: its wheel carries a second HIP runtime)."""
    try:
        import ctypes

        from paraprobe_amd import capi

        lib = capi.load_library()
        h = ctypes.c_void_p()
        if lib.pp_spatstat_create(0, ctypes.byref(h)) != 0:
            return False
        lib.pp_spatstat_destroy(h)
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    """One gfx950 context for the whole GPU test session (fails loudly if the
    native library or the device is missing)."""
    from paraprobe_amd import SpatStat

    c = SpatStat(0)
    yield c
    c.close()
