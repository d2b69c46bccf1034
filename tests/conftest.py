import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqbh.so on cuda:0)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """qbh context on cuda:0 sharing torch's current stream."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from qball_amd import Context
    ctx = Context(0, torch.cuda.current_stream().cuda_stream)
    yield ctx
    ctx.close()
