import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "cmm-turbulence_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def dev():
    """cuda:0 with the HIP library loaded; prints the device once."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import cmm
    print("device:", cmm.device_info(), flush=True)
    return torch.device("cuda:0")
