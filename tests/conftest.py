import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def hip_runtime():
    """PyTorch-ROCm ships its own HIP runtime; in one process it has to reach
    the device before the library's runtime does, or torch finds no GPU.  GPU
    test modules take this fixture before they open a context, so that tests
    handing torch device tensors to the library work in any order."""
    try:
        import torch
    except ImportError:
        return False
    if torch.cuda.is_available():
        torch.cuda.init()
        return True
    return False
