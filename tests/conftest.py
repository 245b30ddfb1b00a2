"""Shared fixtures.  GPU tests are marked ``gpu`` and call the HIP path through
the C ABI; CPU tests cover the oracle, the host logic and the ABI surface."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libct2hip.so")


@pytest.fixture(scope="session")
def orc():
    """The FP64 CPU oracle (test infrastructure)."""
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ct2():
    """The library's Python face, device 0 bound for the whole session."""
    from ct2_amd import api
    api.init(0)
    yield api
    api.finalize()
