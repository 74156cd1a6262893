import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def backend():
    """One backend context per test session (GPU tests only)."""
    import raymotion_amd as rm
    b = rm.Backend(int(os.environ.get("RAYMOTION_HIP_DEVICE", "0")))
    yield b
    b.close()
