import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


@pytest.fixture(scope="session")
def acc():
    from sirius_amd.accelerator import HipAccelerator
    a = HipAccelerator()
    caps = a.getCapabilities()
    print(f"\n[device] {caps['deviceName']} CUs={caps['multiprocessorCount']}", flush=True)
    yield a
    a.close()
