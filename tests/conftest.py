import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def golden():
    from imodel_amd import imxbin

    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = imxbin.read(os.path.join(GOLDEN, name))
        return cache[name]

    return get


def relmax(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rel_l2(a, b, w=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    w = np.ones_like(b) if w is None else w
    return float(np.sqrt(np.sum(w * (a - b) ** 2)) / np.sqrt(np.sum(w * b ** 2)))
