import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def ora():
    """The FP64 CPU oracle (built on first use)."""
    from oracle import oracle as o
    o.build()
    return o


@pytest.fixture(scope="session")
def ctx():
    """One DeviceContext (libdmrg_hip handle on cuda:0) for all GPU tests."""
    from chemtensor_mojo_amd import DeviceContext
    c = DeviceContext(0)
    yield c
    c.close()
