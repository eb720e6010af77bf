"""The ctypes mirrors of the C-ABI structs have the C layout -- CPU.

Compiles a small C program against include/*.h (gcc) that prints sizeof
and offsetof of every field the Python classes declare, and compares them
with ctypes' sizes and offsets: a field added, removed, renamed or moved on
one side only fails here instead of corrupting calls."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = {
    "HostGraphStruct": "hc_host_graph_t",
    "GraphInfo": "hc_graph_info_t",
    "DeviceInfo": "hc_device_info_t",
    "RunStats": "hc_run_stats_t",
    "EpochTrace": "hc_epoch_trace_t",
    "SsspOpts": "hc_sssp_opts_t",
    "BfsOpts": "hc_bfs_opts_t",
    "PrOpts": "hc_pr_opts_t",
    "ScheduleBlock": "hc_schedule_block_t",
    "ServerOpts": "hc_server_opts_t",
    "QueryInfo": "hc_query_info_t",
    "ServerStats": "hc_server_stats_t",
}


def test_ctypes_structs_match_c_headers(hc):
    cc = shutil.which("gcc")
    if not cc:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hc_capi.h"',
             '#include "hc_sched.h"', '#include "hc_batch.h"', '#include "hc_serve.h"',
             "int main(void) {"]
    for py, c in PAIRS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in getattr(hc, py)._fields_:
            lines.append(f'printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as f:
            f.write("\n".join(lines))
        subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), src, "-o", exe],
                       check=True, capture_output=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        py, name, val = line.split()
        got[(py, name)] = int(val)
    for py in PAIRS:
        cls = getattr(hc, py)
        assert got[(py, "size")] == __import__("ctypes").sizeof(cls), py
        for name, _ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, (py, name)
