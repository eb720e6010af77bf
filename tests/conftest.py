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


def _mix64(z):
    """splitmix64's output function (Vigna) on a uint64 array, wrapping."""
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def result_checksum(values) -> int:
    """The library's result checksum restated: the wrapping 64-bit sum over
    v of splitmix64-mix((v << 32) ^ uint32(values[v]))."""
    x = np.ascontiguousarray(values, np.int32).view(np.uint32).astype(np.uint64)
    v = np.arange(x.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(_mix64((v << np.uint64(32)) ^ x).sum(dtype=np.uint64))
