import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librhd.so)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def gpu():
    """One HIP sensor processor for the whole GPU session (librhd.so is required)."""
    from rohditor_amd import GpuSensorProcessor
    proc = GpuSensorProcessor(device=int(os.environ.get("RHD_DEVICE", "0")))
    yield proc
    proc.close()
