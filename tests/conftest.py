import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "yume-pdq_amd"), os.path.join(REPO, "tests"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on cuda:0.  Fails (never skips) when no GPU is usable."""
    import torch
    import yume_pdq
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a usable GPU")
    torch.cuda.init()
    yume_pdq.lib()
    return yume_pdq


@pytest.fixture(scope="session")
def dev():
    import torch
    return torch.device("cuda:0")


@pytest.fixture
def record_cap_4096(gpu):
    """The matrix-core matcher's candidate-record cap lowered to 4096 for one
    test (its overflow path at small sizes), restored afterwards."""
    gpu.set_match_record_cap(4096)
    yield 4096
    gpu.set_match_record_cap(0)
