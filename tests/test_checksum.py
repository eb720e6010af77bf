"""Pins of the result checksum's restatement (conftest.result_checksum) --
CPU.  The device checksum (hc_checksum_i32 / hc_checksum_multi /
hc_server_checksum) is compared with this restatement in the GPU tests."""
from __future__ import annotations

import numpy as np

from conftest import _mix64, result_checksum


def test_mix_is_splitmix64():
    # splitmix64 seeded with 0 adds the golden gamma before mixing, so its
    # first outputs are mix(k * gamma) for k = 0, 1, 2 (published sequence).
    gamma = 0x9E3779B97F4A7C15
    z = np.array([0, gamma, (2 * gamma) % (1 << 64)], np.uint64)
    assert [hex(int(x)) for x in _mix64(z)] == \
        ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4", "0x6c45d188009454f"]


def test_checksum_order_and_position():
    rng = np.random.RandomState(0)
    a = rng.randint(-5, 1 << 30, size=1000).astype(np.int32)
    h = result_checksum(a)
    # a sum over (vertex, value) pairs: permuting values changes it, a single
    # changed value changes it, and splitting the sum adds up
    assert result_checksum(a[::-1]) != h
    b = a.copy()
    b[517] += 1
    assert result_checksum(b) != h
    v = np.arange(a.size, dtype=np.uint64)
    x = a.view(np.uint32).astype(np.uint64)
    parts = _mix64((v << np.uint64(32)) ^ x)
    with np.errstate(over="ignore"):
        assert int(parts[:400].sum(dtype=np.uint64) + parts[400:].sum(dtype=np.uint64)) == h
    assert result_checksum(np.zeros(0, np.int32)) == 0
    # -1 and 0x7fffffff (the unreachable marker) are different values
    assert result_checksum(np.array([-1], np.int32)) != \
        result_checksum(np.array([0x7FFFFFFF], np.int32))
