import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "covof_2d-fortran_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def cv():
    """The product package on the GPU (tests marked gpu only)."""
    import torch
    import covof_amd
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    return covof_amd


def to_gpu(a):
    import torch
    return torch.tensor(a, dtype=torch.float64, device="cuda")
