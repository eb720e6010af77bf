"""Autotuner logic on the CPU: the host/GPU split of a segment profile, the
merged schedule it emits, the cache, and the per-segment host timing."""
from __future__ import annotations

import ctypes as C
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def at(hc):
    return importlib.import_module("hetero-compute_amd.autotune")


def _profile(gpu_ms, cpu_ms, n=1 << 20, alg="sssp"):
    segs = []
    for i, (g, c) in enumerate(zip(gpu_ms, cpu_ms)):
        segs.append({"start": i * 10, "end": i * 10 + 10, "edges": 100, "cpu_ms": c,
                     "gpu": {"SSSP GPU bucketed": (g, 0, 0),
                             "SSSP GPU warp-red": (g * 1.5, 256, 256)}})
    return {"algorithm": alg, "num_nodes": n, "segments": segs}


def test_all_gpu_when_host_is_slow(hc, at):
    blocks, makespan, plan = at.schedule_from_profile(_profile([1.0] * 6, [50.0] * 6))
    assert plan["cpu_segments"] == []
    assert blocks == [{"device": "AMD Instinct MI355X", "start": 0, "end": 5,
                       "kernel": "SSSP GPU bucketed", "blocks": 0, "threads": 0}]
    assert makespan == pytest.approx(6.0)


def test_split_lowers_makespan(hc, at):
    # Segment 2 is slow on the GPU and cheap on the host: it moves; the
    # exchange cost (8 bytes per vertex over 25 GB/s) is charged once.
    gpu = [1.0, 1.0, 9.0, 1.0]
    cpu = [30.0, 30.0, 2.0, 30.0]
    blocks, makespan, plan = at.schedule_from_profile(_profile(gpu, cpu, n=1000))
    assert plan["cpu_segments"] == [2]
    assert makespan == pytest.approx(3.0 + 8000 / 25e9 * 1e3)
    assert [b["device"] for b in blocks] == ["AMD Instinct MI355X", "Host CPU",
                                             "AMD Instinct MI355X"]
    s = hc.Schedule.create(blocks, 4)
    key = lambda b: b["start"]  # noqa: E731  (sections group blocks per device)
    assert sorted(hc.Schedule.parse(s.text()).blocks, key=key) == sorted(s.blocks, key=key)
    # An exchange dearer than the gain keeps everything on the GPU.
    _, _, plan2 = at.schedule_from_profile(_profile(gpu, cpu, n=1 << 32))
    assert plan2["cpu_segments"] == []


def test_best_gpu_kernel_per_segment(hc, at):
    prof = _profile([2.0, 2.0], [100.0, 100.0])
    prof["segments"][1]["gpu"]["SSSP GPU warp-red"] = (0.5, 1024, 256)
    blocks, _, plan = at.schedule_from_profile(prof)
    assert plan["gpu_kernels"] == ["SSSP GPU bucketed", "SSSP GPU warp-red"]
    assert blocks[1] == {"device": "AMD Instinct MI355X", "start": 1, "end": 1,
                         "kernel": "SSSP GPU warp-red", "blocks": 1024, "threads": 256}


def test_cache_roundtrip(hc, at, tmp_path):
    g = hc.HostGraph.generate("kron", 8, 8)
    key = f"dev|sssp|{at.signature(g)}"
    p = str(tmp_path / "c.json")
    at.save(p, key, {"knobs": {"grid_per_cu": 4}})
    at.save(p, "other", {"knobs": {}})
    assert at.load(p, key) == {"knobs": {"grid_per_cu": 4}}
    assert at.load(p, "missing") is None


class _Recorder:
    """Stands in for a DeviceGraph (set_tuning) or a Server
    (set_graph_tuning): records what apply_cached sets."""

    def __init__(self):
        self.calls = []

    def set_tuning(self, key, value):
        self.calls.append((key, value))

    def set_graph_tuning(self, gid, key, value):
        self.calls.append((gid, key, value))


def test_apply_cached_by_signature(hc, at, tmp_path):
    g = hc.HostGraph.generate("kron", 8, 8, weighted=True)
    other = hc.HostGraph.generate("kron", 8, 8, weighted=True, seed=99)
    assert at.signature(g) != at.signature(other)
    p = str(tmp_path / "c.json")
    at.save(p, at.cache_key(g, "sssp"), {"knobs": {"sssp.stage": 8192, "sssp.alpha": 2}})
    at.save(p, at.cache_key(g, "multi"), {"knobs": {"multi.block": 512}})
    dg = _Recorder()
    assert at.apply_cached(p, g, "sssp", dg) == {"sssp.stage": 8192, "sssp.alpha": 2}
    assert sorted(dg.calls) == [("sssp.alpha", 2), ("sssp.stage", 8192)]
    srv = _Recorder()
    assert at.apply_cached(p, g, "multi", (srv, 3)) == {"multi.block": 512}
    assert srv.calls == [(3, "multi.block", 512)]
    # Another graph, another algorithm, or no cache file: nothing applied.
    none = _Recorder()
    assert at.apply_cached(p, other, "sssp", none) == {}
    assert at.apply_cached(p, g, "bfs", none) == {}
    assert at.apply_cached(str(tmp_path / "absent.json"), g, "sssp", none) == {}
    assert none.calls == []


def test_batch_signature(hc, at):
    gs = [hc.HostGraph.generate("kron", 6 + i, 8, seed=i) for i in range(3)]
    s = at.signature(gs)
    assert s.startswith("batch3-") and s == at.signature(list(gs))
    assert s != at.signature(gs[:2]) and s != at.signature(gs[::-1])


def test_many_graphs_workload(hc, at):
    """The autotuner's many-graph workload is the benchmark's: scales
    U{10..16}, edge factors U{8..32}, graph i from seed + 1 + i."""
    gs = at.many_graphs(4, seed=4242)
    rng = np.random.RandomState(4242)
    for i, g in enumerate(gs):
        sc, ef = int(rng.randint(10, 17)), int(rng.randint(8, 33))
        ref = hc.HostGraph.generate("kron", sc, ef, weighted=False, seed=4243 + i)
        assert np.array_equal(g.index, ref.index)
        assert np.array_equal(g.neighbors, ref.neighbors)


@pytest.mark.parametrize("alg", [0, 1, 2])
def test_cpu_epoch_timing(hc, alg):
    g = hc.HostGraph.generate("kron", 12, 16, weighted=True)
    ms = C.c_double()
    hc.check(hc.lib().hc_cpu_epoch_ms(g.struct_ref(), alg, 0, g.num_nodes, 2, C.byref(ms)))
    assert ms.value > 0
    with pytest.raises(hc.HcError):
        hc.check(hc.lib().hc_cpu_epoch_ms(g.struct_ref(), 3, 0, 1, 1, C.byref(ms)))


def test_yaml_profiles_in_reference_layout(hc, at, tmp_path):
    import yaml
    g = hc.HostGraph.generate("kron", 8, 8)
    bounds = g.equal_edge_ranges(3)
    segs = []
    for i in range(3):
        a, b = int(bounds[i]), int(bounds[i + 1])
        segs.append({"start": a, "end": b, "edges": int(g.index[b] - g.index[a]),
                     "cpu_ms": 1.0 + i, "gpu": {"SSSP GPU bucketed": (0.5, 0, 0),
                                                  "SSSP GPU warp-red": (0.7, 256, 256)}})
    files = at.write_yaml_profiles({"algorithm": "sssp", "num_nodes": g.num_nodes,
                                    "segments": segs}, g, str(tmp_path))
    assert len(files) == 3
    docs = [yaml.safe_load(open(f)) for f in files]
    cpu = [d for d in docs if d["config"]["device"] == "Host CPU"][0]
    assert cpu["config"]["kernel"] == "SSSP CPU one-to-one"
    assert [s_["millis"] for s_ in cpu["results"][0]["segments"]] == [1.0, 2.0, 3.0]
    warp = [d for d in docs if d["config"]["kernel"].startswith("SSSP GPU warp-red")][0]
    assert warp["config"]["kernel"] == "SSSP GPU warp-red 256 256"
    # The kernel line of a profile is a valid schedule kernel line.
    assert hc.kernel_id(warp["config"]["kernel"]) == hc.KERNELS.index("SSSP GPU warp-red")
    seg0 = cpu["results"][0]["segments"][0]
    assert seg0["start"] == 0 and seg0["num_edges"] == segs[0]["edges"]


def test_yaml_tree_in_reference_layout(hc, at, tmp_path):
    """Tree profiles follow the reference's tree_res_t writer: one file per
    device and kernel, `results` a list of `{depth, segments}` with 2^depth
    equal-edge segments per layer."""
    import yaml
    g = hc.HostGraph.generate("kron", 8, 8)
    layers = []
    for depth in range(3):
        bounds = g.equal_edge_ranges(1 << depth)
        segs = []
        for i in range(1 << depth):
            a, b = int(bounds[i]), int(bounds[i + 1])
            segs.append({"start": a, "end": b, "edges": int(g.index[b] - g.index[a]),
                         "cpu_ms": 1.0 + i, "gpu": {"BFS GPU bucketed": (0.25, 0, 0)}})
        layers.append({"algorithm": "bfs", "num_nodes": g.num_nodes, "segments": segs})
    files = at.write_yaml_tree({"algorithm": "bfs", "num_nodes": g.num_nodes,
                                "layers": layers}, g, str(tmp_path))
    assert len(files) == 2
    docs = {yaml.safe_load(open(f))["config"]["device"]: yaml.safe_load(open(f)) for f in files}
    cpu = docs["Host CPU"]
    assert [r["depth"] for r in cpu["results"]] == [0, 1, 2]
    for r in cpu["results"]:
        segs = r["segments"]
        assert len(segs) == 1 << r["depth"]
        assert segs[0]["start"] == 0 and segs[-1]["end"] == g.num_nodes
        assert sum(x["num_edges"] for x in segs) == g.num_edges
    gpu = docs["AMD Instinct MI355X"]
    assert hc.kernel_id(gpu["config"]["kernel"]) == hc.KERNELS.index("BFS GPU bucketed")
