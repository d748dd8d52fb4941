import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def gpu_ready():
    import acvd_amd
    n = acvd_amd.lib().acvd_device_count()
    if n < 1:
        pytest.fail("GPU test collected but no HIP device is visible")
    return True
