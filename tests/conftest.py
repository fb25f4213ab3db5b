import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def geometry():
    from allen_amd import synth
    return synth.read_geometry()


@pytest.fixture(scope="session")
def beamline():
    from allen_amd import synth
    return synth.read_beamline()
