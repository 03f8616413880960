import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def ckd():
    """The GPU package (skips when no HIP device is visible)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import cosmokdtree_amd
    return cosmokdtree_amd
