import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libphx.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context for the whole GPU session (one process, one GPU)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from photon_amd.backend import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
