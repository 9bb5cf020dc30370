import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built native library")


@pytest.fixture(scope="session")
def prover():
    """One device context for all GPU tests of the session."""
    from circlecover_amd import Prover

    p = Prover(0)
    yield p
    p.close()
