import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libprism_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: CPU test that takes a minute or more")


@pytest.fixture(scope="session")
def ref():
    """FP64 CPU oracle backend (checker)."""
    import oracle
    return oracle.backend()


@pytest.fixture(scope="session")
def hip():
    """HIP backend on device 0; fails loudly when the library or the GPU is missing."""
    import prism_amd as P
    be = P.Backend.hip(0)
    yield be
    be.close()
