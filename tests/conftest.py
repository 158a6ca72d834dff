import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


def gpu_available() -> bool:
    try:
        from svfsi_amd import fsils

        return fsils.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly (no skip) when a GPU test runs without a HIP device or
    without the built library: the HIP path must be the one that runs."""
    from svfsi_amd import fsils

    n = fsils.device_count()
    assert n > 0, "gpu test needs a HIP device and libsvfsi_amd.so"
    import torch

    assert torch.cuda.is_available()
    return torch.device("cuda")
