import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref")
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


@pytest.fixture(autouse=True)
def _free_gpu_objects():
    """Release the GPU handles of a test (and of earlier failed tests) in the main thread at a known
    point, instead of whenever the garbage collector runs."""
    yield
    import gc
    gc.collect()


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


def load_sw():
    """Import the package from shallow-water_amd/ (its directory name is not an identifier)."""
    import importlib.util
    if "shallow_water_amd" in sys.modules:
        return sys.modules["shallow_water_amd"]
    pkg = os.path.join(ROOT, "shallow-water_amd")
    spec = importlib.util.spec_from_file_location("shallow_water_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["shallow_water_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def sw():
    return load_sw()


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)
