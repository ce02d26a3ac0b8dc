import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def gpu():
    """The initialized native library; GPU tests fail loudly if the HIP path cannot run."""
    import warpstore_amd as ws
    L = ws.lib()
    rc = L.ws_init()
    assert rc == 0, f"GPU init failed: {L.ws_last_error().decode()}"
    return L
