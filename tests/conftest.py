import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU-only check")


def _gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU fixture: a gpu-marked test on a box without a GPU is an error, not a pass."""
    if not _gpu_available():
        pytest.fail("gpu-marked test needs a GPU (torch.cuda.is_available() is False)")
    from bine_trees_fugaku_amd import _lib
    return _lib.load()
