import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def ctx():
    """One library context on device 0 for the whole GPU session."""
    from crysplight_amd import _lib
    return _lib.default_context(0)
