"""Build checks on the compiler's per-kernel resource report (written next to
each object by the library's Makefile): no kernel spills registers to scratch
memory, and the loops' step kernels keep enough wavefronts per SIMD to hide
gather latency."""
from __future__ import annotations

import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "hetero-compute_amd", "_build")


def _reports():
    files = sorted(glob.glob(os.path.join(BUILD, "*.hip.o.res")))
    if not files:
        pytest.skip("library not built here (no resource reports)")
    kernels = {}
    for f in files:
        cur = None
        for line in open(f):
            m = re.search(r"Function Name: (\S+)", line)
            if m:
                cur = m.group(1)
                kernels[cur] = {}
                continue
            m = re.search(r"remark:\s+(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)",
                          line)
            if m and cur:
                kernels[cur][m.group(1).split()[0]] = int(m.group(2))
    return kernels


def test_no_kernel_spills():
    kernels = _reports()
    assert len(kernels) >= 20
    spills = {k: v["ScratchSize"] for k, v in kernels.items() if v.get("ScratchSize", 0)}
    assert not spills, spills


def test_step_kernels_occupancy():
    kernels = _reports()
    steps = {k: v for k, v in kernels.items() if "step_kernel" in k}
    assert steps
    low = {k: v["Occupancy"] for k, v in steps.items() if v["Occupancy"] < 5}
    assert not low, low
