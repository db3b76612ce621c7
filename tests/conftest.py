import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE = os.path.join(ROOT, "oracle")
if ORACLE not in sys.path:
    sys.path.insert(0, ORACLE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def announce(msg):
    """Name the case before a GPU step, flushed, so a lost box still says
    what was running."""
    print("[gpu-case] " + msg, flush=True)


@pytest.fixture(scope="session")
def oracle_glibc():
    import pyoracle
    return pyoracle.Oracle(bxpow=False)


@pytest.fixture(scope="session")
def oracle_bxpow():
    import pyoracle
    return pyoracle.Oracle(bxpow=True)
