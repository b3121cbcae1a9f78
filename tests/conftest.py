import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_collection_finish(session):
    """Some GPU tests hold device buffers in torch tensors. torch carries its
    own HIP runtime, which finds no GPU once the library's runtime has opened
    it in the same process, so torch opens it first whenever GPU tests run."""
    if not any(item.get_closest_marker("gpu") for item in session.items):
        return
    import torch
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda:0")
        torch.cuda.synchronize()


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    return oracle
