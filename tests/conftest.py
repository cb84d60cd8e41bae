import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # Some GPU tests allocate device buffers with torch.  torch ships its own
    # HIP runtime: it must initialise the device before libsbamd.so loads the
    # system one, else torch finds no GPU.  Only when GPU tests are selected.
    expr = config.getoption("markexpr", "") or ""
    if "gpu" in expr and "not gpu" not in expr:
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except ImportError:
            pass


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")
