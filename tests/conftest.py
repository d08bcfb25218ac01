import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built libjoker.so")


@pytest.fixture(scope="session")
def gpu():
    """Initialise libjoker on device 0.  Fails (does not skip) when the
    library or the device is missing: GPU tests must run the HIP path."""
    import torch

    from thejoker_amd import _lib
    assert torch.cuda.is_available(), "gpu-marked test run without a visible GPU"
    _lib.ensure_device(0)
    return torch
