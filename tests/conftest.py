"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "black-hole_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def bt():
    import bhtrace
    bhtrace.load()
    return bhtrace


@pytest.fixture(scope="session")
def oracle(bt):
    """The CPU oracle, with its FP64 band rule set like the library's default
    precision (tests that change one change the other)."""
    import pyoracle
    O = pyoracle.Oracle()
    p = bt.get_precision()
    O.set_band(p.fp64Band, p.critBand, p.polarBand)
    return O


@pytest.fixture(scope="session")
def gpu(bt):
    if bt.load().bhDeviceCount() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return bt
