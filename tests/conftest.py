"""Shared fixtures: the package (via its directory name), the oracle, and
small graphs.  Tests marked `gpu` need an MI355X; everything else runs on
the CPU."""
from __future__ import annotations

import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X)")


@pytest.fixture(scope="session")
def hc():
    return importlib.import_module("hetero-compute_amd")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def gpu(hc):
    """Initialises GPU 0 (raises if the library or the device is missing)."""
    hc.check(hc.lib().hc_init(0))
    return hc.device_info()


def golden_files():
    return sorted(f for f in os.listdir(GOLDEN_DIR) if f.endswith(".npz"))


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN_DIR, name), allow_pickle=False))
