import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libfsx.so")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Session guard for GPU tests: the GPU and the native library must both be there
    (a missing library is a failure, not a skip: there is no fallback path)."""
    if not gpu_available():
        pytest.skip("no ROCm GPU visible")
    from fieldstitch_amd import _lib
    _lib.load()
    return True
