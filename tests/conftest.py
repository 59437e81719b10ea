import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O._load()
    return O


@pytest.fixture(scope="session")
def gpu():
    """Torch device handle; fails loudly if the HIP library or the GPU is missing."""
    import torch
    import lifted_amd
    from lifted_amd import _lib
    _lib.lib()
    if not torch.cuda.is_available():
        pytest.fail("GPU test requested but no ROCm device is visible")
    return torch.device("cuda:0")
