import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def kats():
    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    """The product package with its HIP library loaded and a device present."""
    import mpsft_amd
    mpsft_amd.lib()
    if mpsft_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return mpsft_amd


def approx(a, b, eps=1.1920928955078125e-05):
    """Catch 1.9.6 Approx (src/catch.hpp:2842-2875)."""
    return abs(a - b) < eps * (1.0 + max(abs(a), abs(b)))
