import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libmyql_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def engine():
    """The HIP engine on device 0. Fails (never skips) when the library or the
    device is missing: GPU tests must run the native path."""
    from myquantlib_amd.pricing import Engine
    eng = Engine(0)
    yield eng
    eng.close()
