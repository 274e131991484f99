import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running check at full size")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The HIP library and torch on cuda:0; fails (does not skip) without a GPU."""
    import torch
    assert torch.cuda.is_available(), "gpu test run without a visible GPU"
    import mitos_amd.ic as ic
    ic.lib()
    return ic
