import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def orc():
    from oracle.oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def orc_libm():
    from oracle.oracle import Oracle

    return Oracle(libm=True)


@pytest.fixture(scope="session")
def u1():
    """The gfx950 library, initialised on device 0 (GPU tests only)."""
    import latticeu1_amd as u1

    u1.init(0)
    print(f"[u1] device: {u1.device_name()}", flush=True)
    yield u1
