import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through libgcq's C ABI)")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    pyoracle.load()
    return pyoracle


@pytest.fixture(scope="session")
def gcq():
    import gcq_amd
    gcq_amd.load()  # raises if libgcq.so is missing: there is no fallback
    return gcq_amd


GOLDEN = os.path.join(ROOT, "tests", "golden")
