import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not os.path.exists(oracle.LIB):
        oracle.build()
    return oracle.lib()


@pytest.fixture(scope="session")
def gpu():
    """Loads libpbpic.so (built beforehand) on a GPU box; fails loudly when the
    library or the GPU is missing."""
    import pbpic_amd
    if not os.path.exists(pbpic_amd.LIB_PATH):
        from pbpic_amd import build
        build.build_library()
    import torch
    assert torch.cuda.is_available(), "gpu tests need a GPU"
    return pbpic_amd.load_library()
