import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def backend():
    """One HIP device context shared by the GPU tests."""
    from fastwarp_amd.backend import HipBackend

    b = HipBackend(int(os.environ.get("FW_DEVICE", "0")))
    yield b
    b.close()
