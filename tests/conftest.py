import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU")


@pytest.fixture(scope="session")
def cpu_oracle():
    import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build()
    return oracle.CpuRenderer("oracle")


@pytest.fixture(scope="session")
def cpu_ref():
    import oracle
    if not oracle.ref_available() and os.path.isdir(oracle.REF_SRC):
        oracle.build()
    if not oracle.ref_available():
        pytest.skip("the reference build (oracle/_ref) is not available here")
    return oracle.CpuRenderer("ref")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
