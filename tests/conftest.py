import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "kinectfusion-cuda_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C ABI")


@pytest.fixture(scope="session")
def oracle32():
    from oracle.kfref import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle64():
    import numpy as np
    from oracle.kfref import Oracle
    return Oracle(np.float64)
