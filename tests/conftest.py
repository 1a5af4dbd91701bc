import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libiac_hip.so")


@pytest.fixture(scope="session")
def dev():
    import torch

    assert torch.cuda.is_available(), "GPU test collected on a host without a HIP device"
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def ctx(dev):
    import incohautocorrelate_amd as iac

    c = iac.Context(0)
    yield c
    c.close()
