import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    """Tests load what is built; they build only when nothing is."""
    from partapp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from partapp_amd import build
        build.build()
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
