import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libflx on cuda:0)")


@pytest.fixture(scope="session")
def gpu():
    """The device path: libflx.so loaded and a GPU present.  No CPU fallback:
    a GPU test run without either fails instead of skipping."""
    import torch
    import flens_amd as F
    F.lib()
    if not torch.cuda.is_available():
        raise RuntimeError("GPU test run without a GPU")
    torch.cuda.init()
    return F
