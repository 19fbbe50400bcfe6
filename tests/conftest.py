"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs on
the CPU.  The oracle (oracle/pyoracle.py) is imported here to check the HIP results."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def orc():
    import pyoracle

    if not pyoracle.LIB_PATH.exists():
        pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gs():
    import gpusteer_amd

    return gpusteer_amd


@pytest.fixture(scope="session")
def ctx(gs):
    c = gs.Context(0)
    yield c
    c.close()
