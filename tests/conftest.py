import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libccr_hip.so")


@pytest.fixture(scope="session")
def engine():
    from zlatko_amd import native
    e = native.Engine()
    yield e
    e.close()
