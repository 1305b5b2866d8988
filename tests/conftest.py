import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libticc_hip.so")


@pytest.fixture(scope="session")
def gpu_device():
    import torch

    from pyticc_amd import _device

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a HIP device")
    return _device.require_gpu(0)
