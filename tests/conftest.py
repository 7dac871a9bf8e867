import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """Initialise the HIP library on device 0; fails loudly without it."""
    import torch
    from falm_amd import runtime as rt
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    rt.lib()
    rt.init(0)
    torch.cuda.set_device(0)
    return rt
