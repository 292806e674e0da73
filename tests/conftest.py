import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the product kernels")


@pytest.fixture(scope="session")
def gpu():
    """libqpms_hip with an initialised device (GPU tests only)."""
    import qpms_amd
    L = qpms_amd.lib()
    rc = L.qpms_hip_init()
    assert rc == 0, "no HIP device: GPU tests need an MI355X"
    return L
