import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

TABLES = os.path.join(ROOT, "specter_amd", "tables")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libspx.so")


@pytest.fixture(scope="session")
def tables_dir():
    return TABLES


def rel_err(a, b):
    """max|a-b| / max|b| (0 if both are zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0:
        return float(num)
    return float(num / den)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den else 1.0))
