"""Segment schedules on the CPU: .skd parsing and writing, the reference's
kernel and device names, schedule validation, and the co-execution runtime
driven by schedules whose blocks are all host-core blocks (no GPU needed).

Reference: src/schedule.h:32-226 (types, get_kernel_id, load_sssp_schedule),
util/scheduler/sched.py:263-286 (write_schedule), src/devices.h:16-90,
kernel name strings in src/kernels/{cpu,gpu}/{sssp,pr,bfs}.cuh."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden

PR_L1_TOL = 1e-6

# The segment split of the reference's generated SSSP driver
# (src/kernels/heterogeneous/sssp.cuh:43-72, 281-290): 24 segments, GPU
# blocks over runs of segments, segments 17 and 22 on the host.
REFERENCE_LAYOUT = """NVIDIA Quadro RTX 4000
0 0
SSSP GPU block-red 256 1024
1 1
SSSP GPU block-red 1024 256
2 2
SSSP GPU block-red 512 512
3 4
SSSP GPU block-red 256 1024
5 5
SSSP GPU warp-red 256 1024
6 9
SSSP GPU warp-red 256 1024
10 14
SSSP GPU warp-red 256 1024
15 16
SSSP GPU warp-red 256 1024
18 18
SSSP GPU warp-red 256 1024
19 21
SSSP GPU warp-red 256 1024
23 23
SSSP GPU one-to-one 256 1024
Intel i7-9700K
17 17
SSSP CPU one-to-one
22 22
SSSP CPU one-to-one
"""


def test_names_roundtrip(hc):
    for i, name in enumerate(hc.KERNELS):
        assert hc.kernel_id(name) == i
        assert hc.lib().hc_kernel_to_string(i).decode() == name
    for i, name in enumerate(hc.DEVICES):
        assert hc.device_id(name) == i
        assert hc.lib().hc_device_to_string(i).decode() == name
    # get_kernel_id matches by prefix (src/schedule.h:133-142).
    assert hc.kernel_id("SSSP GPU block-red 256 1024") == hc.KERNELS.index("SSSP GPU block-red")
    with pytest.raises(hc.HcError):
        hc.kernel_id("SSSP GPU block")       # a prefix of a name is not a name
    with pytest.raises(hc.HcError):
        hc.device_id("NVIDIA Quadro")         # devices match whole lines only


def test_parse_reference_layout(hc):
    s = hc.Schedule.parse(REFERENCE_LAYOUT)
    assert s.num_segments == 24
    b = s.blocks
    assert len(b) == 13
    assert b[0] == {"device": "NVIDIA Quadro RTX 4000", "device_id": 0, "start": 0,
                    "end": 0, "kernel": "SSSP GPU block-red", "blocks": 256,
                    "threads": 1024}
    assert b[1]["blocks"] == 1024 and b[1]["threads"] == 256
    assert [x["kernel"] for x in b[-2:]] == ["SSSP CPU one-to-one"] * 2
    assert all(x["device"] == "Intel i7-9700K" for x in b[-2:])
    # Writer emits the same text back (write_schedule layout).
    assert s.text() == REFERENCE_LAYOUT
    assert hc.Schedule.parse(s.text()).blocks == b


def test_device_instances_and_crlf(hc, tmp_path):
    text = ("AMD Instinct MI355X\r\n0 1\r\nSSSP GPU bucketed\r\n"
            "Host CPU\n2 2\nSSSP CPU one-to-one\n\n"
            "AMD Instinct MI355X\n3 3\nSSSP GPU warp-red 128 256\n")
    p = tmp_path / "x.skd"
    p.write_bytes(text.encode())
    s = hc.Schedule.load(str(p))
    b = s.blocks
    assert s.num_segments == 4
    assert [(x["device"], x["device_id"]) for x in b] == [
        ("AMD Instinct MI355X", 0), ("Host CPU", 0), ("AMD Instinct MI355X", 1)]
    assert b[0]["blocks"] == 0 and b[0]["threads"] == 0  # library default
    assert b[2]["blocks"] == 128 and b[2]["threads"] == 256


@pytest.mark.parametrize("text,msg", [
    ("0 1\nSSSP CPU one-to-one\n", "neither a device"),
    ("Host CPU\n0 1\n", "without a kernel"),
    ("Host CPU\n0\nSSSP CPU one-to-one\n", "expected"),
    ("Host CPU\n2 1\nSSSP CPU one-to-one\n", "expected"),
    ("Host CPU\n0 1\nSSSP GPU warp-red 1 64\n", "cannot run"),
    ("AMD Instinct MI355X\n0 1\nSSSP CPU one-to-one\n", "cannot run"),
    ("AMD Instinct MI355X\n0 1\nSSSP GPU warp-red 64\n", "expected"),
    ("AMD Instinct MI355X\n0 1\nSSSP GPU warp-red 64 2048\n", "expected"),
    ("Host CPU\n0 1\nSSSP CPU one-to-one 4 4\n", "no launch shape"),
    ("Host CPU\n0 1\nSSSP Quantum\n", "unknown kernel"),
    ("", "no blocks"),
])
def test_parse_errors(hc, text, msg):
    with pytest.raises(hc.HcError) as e:
        hc.Schedule.parse(text)
    assert e.value.code == -1 and msg in str(e.value)


def test_load_missing_file(hc, tmp_path):
    with pytest.raises(hc.HcError) as e:
        hc.Schedule.load(str(tmp_path / "none.skd"))
    assert e.value.code == -4


def _fixture_graph(hc, name):
    fx = load_golden(name)
    return fx, hc.HostGraph.from_csr(fx["index"], fx["neighbors"], fx["weights"])


def _cpu_schedule(hc, algo, parts, segments):
    kern = {"sssp": "SSSP CPU one-to-one", "pr": "PR CPU one-to-one",
            "bfs": "BFS CPU pull"}[algo]
    blocks = [{"device": "Host CPU" if i % 2 == 0 else "Intel Xeon E5-2686",
               "start": a, "end": b, "kernel": kern} for i, (a, b) in enumerate(parts)]
    return hc.Schedule.create(blocks, segments)


def test_coverage_validation(hc):
    fx, g = _fixture_graph(hc, "kron8.npz")
    with pytest.raises(hc.HcError) as e:
        hc.Hetero(g, _cpu_schedule(hc, "sssp", [(0, 2), (4, 5)], 6))
    assert "no block" in str(e.value)
    with pytest.raises(hc.HcError) as e:
        hc.Hetero(g, _cpu_schedule(hc, "sssp", [(0, 3), (3, 5)], 6))
    assert "blocks 0 and 1" in str(e.value)
    mixed = hc.Schedule.create(
        [{"device": "Host CPU", "start": 0, "end": 0, "kernel": "SSSP CPU one-to-one"},
         {"device": "Host CPU", "start": 1, "end": 1, "kernel": "PR CPU one-to-one"}], 2)
    with pytest.raises(hc.HcError) as e:
        hc.Hetero(g, mixed)
    assert "another algorithm" in str(e.value)
    with pytest.raises(hc.HcError):
        hc.Schedule.create([{"device": "Host CPU", "start": 0, "end": 3,
                             "kernel": "SSSP CPU one-to-one"}], 3)


@pytest.mark.parametrize("name", ["kron8.npz", "uniform8.npz", "kron10.npz",
                                  "disconnected.npz", "path10.npz"])
def test_cpu_schedule_sssp(hc, name):
    fx, g = _fixture_graph(hc, name)
    with hc.Hetero(g, _cpu_schedule(hc, "sssp", [(0, 1), (2, 4), (5, 5)], 6)) as h:
        for i, s in enumerate(fx["sources"]):
            r = h.sssp(int(s))
            assert np.array_equal(r.values, fx["sssp"][i]), (name, s)
            assert r.stats["epochs"] >= 1


@pytest.mark.parametrize("name", ["kron8.npz", "kron10.npz", "star9.npz", "cycle12.npz"])
def test_cpu_schedule_bfs(hc, oracle, name):
    fx, g = _fixture_graph(hc, name)
    with hc.Hetero(g, _cpu_schedule(hc, "bfs", [(0, 3), (4, 7)], 8)) as h:
        for i, s in enumerate(fx["sources"]):
            s = int(s)
            r = h.bfs(s)
            assert np.array_equal(r.values, fx["depth"][i]), (name, s)
            assert oracle.verify_parents(fx["depth"][i], r.parents, s) == 0
            _, _, dirs = oracle.bfs_do(fx["index"], fx["neighbors"], s)
            assert r.stats["pull_epochs"] == int((dirs == 1).sum())
            assert r.stats["push_epochs"] == int((dirs == 0).sum())


@pytest.mark.parametrize("name", ["kron8.npz", "kron10.npz", "complete6.npz"])
def test_cpu_schedule_pagerank(hc, oracle, name):
    fx, g = _fixture_graph(hc, name)
    with hc.Hetero(g, _cpu_schedule(hc, "pr", [(0, 0), (1, 3)], 4)) as h:
        r = h.pagerank(tolerance=-1.0, max_iters=20)
    assert r.stats["epochs"] == 20
    ref = oracle.pr_jacobi_f32(fx["index"], fx["neighbors"], 0.85, 20)
    assert np.abs(r.values.astype(np.float64) - ref).sum() <= PR_L1_TOL
    assert np.abs(r.values.astype(np.float64) - fx["pr20"]).sum() <= PR_L1_TOL


def test_cpu_sssp_epoch_export(hc, oracle):
    """hc_cpu_sssp_pull_epoch called by a one-thread 'team' is a Gauss-Seidel
    sweep in vertex order; repeated to a fixpoint it gives Dijkstra."""
    import ctypes as C
    fx, g = _fixture_graph(hc, "kron8.npz")
    s = int(fx["sources"][0])
    d = np.full(g.num_nodes, hc.INF_DIST, np.int32)
    d[s] = 0
    upd = np.zeros(1, np.int32)
    for _ in range(g.num_nodes + 1):
        upd[0] = 0
        hc.lib().hc_cpu_sssp_pull_epoch(g.struct_ref(), d.ctypes.data_as(C.c_void_p), 0,
                                        g.num_nodes, 0, 1, upd.ctypes.data_as(C.c_void_p))
        if upd[0] == 0:
            break
    assert np.array_equal(d, fx["sssp"][0])


def test_bfs_push_kernel_names_parse_and_run(hc, oracle):
    """The push-epoch kernel names the reference's push_pull_scheduler emits
    ("BFS CPU push by node" / "... by edge") parse as CPU kernels, cover no
    segment, and choose how host top-down epochs split their work; depths are
    the same either way, and a schedule of push blocks only pulls on the host
    cores."""
    fx, g = _fixture_graph(hc, "kron10.npz")
    for push in ("BFS CPU push by node", "BFS CPU push by edge"):
        s = hc.Schedule.parse(f"Host CPU\n0 3\nBFS CPU pull\nIntel i7-9700K\n0 3\n{push}\n")
        blocks = s.blocks
        assert [b["kernel"] for b in blocks] == ["BFS CPU pull", push]
        for sched in (s, hc.Schedule.parse(f"Host CPU\n0 1\n{push}\n")):
            with hc.Hetero(g, sched) as h:
                for i, src in enumerate(fx["sources"]):
                    src = int(src)
                    r = h.bfs(src)
                    assert np.array_equal(r.values, fx["depth"][i]), (push, src)
                    assert oracle.verify_parents(fx["depth"][i], r.parents, src) == 0
                    _, _, dirs = oracle.bfs_do(fx["index"], fx["neighbors"], src)
                    assert r.stats["push_epochs"] == int((dirs == 0).sum())
    with pytest.raises(hc.HcError) as e:
        hc.Schedule.parse("AMD Instinct MI355X\n0 1\nBFS CPU push by node\n")
    assert "cannot run" in str(e.value)
