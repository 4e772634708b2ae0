import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for d in (os.path.join(ROOT, "exafmm-t_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(__file__)):
    if d not in sys.path:
        sys.path.insert(0, d)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    # The engine links the system ROCm runtime; torch bundles its own copy and
    # the first runtime initialised in a process owns the device, so GPU
    # detection goes through the engine library and tests never import torch.
    from exafmm import abi
    return abi.device_count() > 0


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.skip("no GPU")
