import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def qmc_of(z):
    """[QMC] dictionary of a golden fixture's parameter string."""
    q = {}
    for line in str(z["param_string"]).split("\n"):
        if "=" in line:
            k, v = line.split("#")[0].split("=", 1)
            q[k.strip()] = v.strip()
    return q


@pytest.fixture
def golden():
    return load_golden
