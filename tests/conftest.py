import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")


@pytest.fixture(scope="session")
def lv():
    from __graft_entry__ import load_package

    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import lvoracle

    return lvoracle


@pytest.fixture(scope="session")
def gpu():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    torch.cuda.set_device(0)
    return torch.cuda.get_device_name(0)
