import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ref():
    """The reference Fortran oracle (oracle/_ref, built by `make -C oracle`)."""
    from oracle import ref_shdom
    if not ref_shdom.available():
        pytest.skip("reference oracle library not built (make -C oracle)")
    return ref_shdom


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
