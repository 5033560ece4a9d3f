import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpsx.so")
    config.addinivalue_line("markers", "slow: long-running test")


@pytest.fixture(scope="session")
def psx():
    """The product bindings; GPU tests skip only when no HIP device is visible."""
    import peasoup_amd as P
    P.lib()  # raises if the library is not built
    if not P.gpu_available():
        pytest.skip("no HIP device visible")
    return P


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def psx_host():
    """The product bindings for host-only entry points (no device needed)."""
    import peasoup_amd as P
    P.lib()  # raises if the library is not built
    return P
