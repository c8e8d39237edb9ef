import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def gpu():
    """Selects device 0; GPU tests fail loudly when no device is visible."""
    from belewitte_amd import device
    n = device.device_count()
    assert n >= 1, "GPU test needs a visible GPU"
    device.set_device(0)
    return device.device_props(0)
