import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libctg.so on cuda:0)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc
    orc.load()
    return orc


@pytest.fixture(scope="session")
def ctg_lib():
    import ct_amd
    return ct_amd.load()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU (run with -m 'not gpu' on CPU hosts)")
    return torch.device("cuda:0")
