import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def ref_element():
    return golden("ref_element.npz")


@pytest.fixture(scope="session")
def ref_assembly():
    return golden("ref_assembly.npz")


@pytest.fixture(scope="session")
def ref_be_cube10():
    return golden("ref_be_cube10.npz")


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max() if b.size else 1.0
    return float(np.abs(a - b).max() / (scale if scale > 0 else 1.0)) if b.size else 0.0


def per_item_rel_err(a, b, axis_items=0):
    """max over items of (max |a_i - b_i| / max |b_i|): each element's block measured on its own scale."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    a2 = a.reshape(a.shape[0], -1)
    b2 = b.reshape(b.shape[0], -1)
    scale = np.maximum(np.abs(b2).max(axis=1), np.finfo(float).tiny)
    return float((np.abs(a2 - b2).max(axis=1) / scale).max()) if b2.size else 0.0
