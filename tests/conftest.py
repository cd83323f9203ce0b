import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C ABI")


@pytest.fixture(scope="session")
def ref_dumps():
    import numpy as np
    return np.load(os.path.join(GOLDEN, "reference_dumps.npz"))


@pytest.fixture(scope="session")
def ctx():
    import largetb_amd as L
    c = L.Context(0)
    yield c
    c.close()


def dense_from_triplets(d, name):
    import numpy as np
    sh = tuple(d[name + "_shape"])
    A = np.zeros(sh)
    A[d[name + "_row"], d[name + "_col"]] = d[name + "_val"]
    return A
