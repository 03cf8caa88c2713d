import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import picref
    picref.build()
    return picref


def gpu_available():
    try:
        from picfoam_amd import picmi
        return picmi.device_count() > 0
    except Exception:
        return False
