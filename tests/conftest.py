import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP kernels")
    config.addinivalue_line("markers", "slow: long physics validation")


@pytest.fixture(scope="session")
def gpu_lib():
    """The loaded HIP library (fails loudly when it is missing)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    from vivsim_amd import _lib
    return _lib.lib()
