import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def be64():
    import numpy as np
    import phist_amd
    be = phist_amd.Backend(np.float64)
    be.init()
    return be


@pytest.fixture(scope="session")
def be32():
    import numpy as np
    import phist_amd
    be = phist_amd.Backend(np.float32)
    be.init()
    return be
