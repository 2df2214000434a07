import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref")
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the HIP library")


def gold(rel):
    """Load one reference fixture (big-endian f64, Fortran order, no header) as a flat array."""
    return np.fromfile(os.path.join(GOLD, rel), dtype=">f8").astype(np.float64)


def gold_u64(rel):
    return np.fromfile(os.path.join(GOLD, rel), dtype=">u8").astype(np.uint64)


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    _oracle.lib()
    return _oracle
