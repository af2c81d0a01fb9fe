import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
# the reference's v5 checkpoint (stereonet_weights.txt), shipped with the package
V5 = os.path.join(ROOT, "from-scratch-stereo-vision_amd", "weights", "stereonet_weights_v5.txt")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def package():
    """The product package (its directory name is not a Python identifier)."""
    return importlib.import_module("from-scratch-stereo-vision_amd")


@pytest.fixture(scope="session")
def sv():
    return package()
