import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lcurve-rs_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
# CPU setup threads: the CPU share of one GPU on a shared node.
os.environ.setdefault("LCURVE_HIP_THREADS", "16")
for p in (PKG, ORACLE, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the engine through the C ABI")


@pytest.fixture(scope="session")
def golden_model_text():
    with open(os.path.join(GOLDEN, "test_model.dat")) as f:
        return f.read()


@pytest.fixture(scope="session")
def golden_lc():
    import numpy as np
    return np.loadtxt(os.path.join(GOLDEN, "test_lc.dat"))


@pytest.fixture(scope="session")
def lcurve_rs():
    import lcurve_rs as lc
    if not lc.has_hip():
        pytest.fail("no HIP device visible: GPU tests need the engine on a GPU (no CPU fallback)")
    lc.set_device("hip:0")
    return lc
