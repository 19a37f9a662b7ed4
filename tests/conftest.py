import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libxdem_hip.so)")


@pytest.fixture(scope="session")
def gpu_available():
    import torch  # noqa: F401  (device discovery only)
    return torch.cuda.is_available()
