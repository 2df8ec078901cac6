import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcadr_hip.so on cuda:0)")


def _gpu_present() -> bool:
    try:
        from cadr_amd import device
        return device.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Selects device 0; GPU tests fail loudly (no fallback) if the library or device is missing."""
    from cadr_amd import device
    n = device.device_count()
    assert n > 0, "no HIP device visible"
    device.set_device(0)
    return 0


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle
