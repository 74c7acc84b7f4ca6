import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "high-order-regularization_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 (the parity metric; pointwise relative error is not used)."""
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def rel_max(a, b):
    """max|a - b| / max|b|."""
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref2():
    from pyoracle import Ref2
    path = os.path.join(ORACLE, "_ref", "libref2surf.so")
    if not os.path.exists(path):
        pytest.skip("reference build (oracle/_ref) not present")
    return Ref2()


@pytest.fixture(scope="session")
def refgrid():
    from pyoracle import RefGrid
    path = os.path.join(ORACLE, "_ref", "librefgrid.so")
    if not os.path.exists(path):
        pytest.skip("reference build (oracle/_ref) not present")
    return RefGrid()


@pytest.fixture(scope="session")
def ctx():
    import hor
    c = hor.Context(0)
    yield c
    c.close()


def golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} missing")
    return np.load(path, allow_pickle=False)
