import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device; runs the HIP library")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def _gpu_available() -> bool:
    try:
        from liquidsimulation_amd.solver import device_info
        name, _ = device_info()
        return name.startswith("gfx950")
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIBORC):
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """Skips unless the gfx950 library can create a context.  The library itself
    must be built: a missing liblps.so fails loudly instead of skipping."""
    from liquidsimulation_amd import _lib
    _lib.lib()  # raises LibraryMissing
    if not _gpu_available():
        pytest.skip("no gfx950 device")
    return True
