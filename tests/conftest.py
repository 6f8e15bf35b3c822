import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    # A gpu-marked test that runs without a GPU fails rather than skips, so a
    # GPU run can never pass by skipping.
    if not has_gpu():
        pytest.fail("gpu-marked test needs a GPU (run CPU tests with -m 'not gpu')")
    import torch
    return torch.device("cuda:0")
