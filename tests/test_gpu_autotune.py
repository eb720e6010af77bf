"""GPU tests of the autotuner: knob search keeps results exact and applies
the best setting; a measured segment schedule runs bit-exact through the
co-execution runtime."""
from __future__ import annotations

import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def at(hc):
    return importlib.import_module("hetero-compute_amd.autotune")


@pytest.fixture(scope="module")
def kron14(hc):
    return hc.HostGraph.generate("kron", 14, 16, weighted=True)


@pytest.mark.parametrize("algorithm", ["sssp", "bfs", "multi"])
def test_tune_knobs(hc, gpu, at, kron14, algorithm):
    before = hc.tuning()
    space = {k: v[:2] for k, v in list(at.KNOB_SPACE[algorithm].items())[:3]}
    sources = kron14.pick_sources(64 if algorithm == "multi" else 2)
    try:
        with hc.DeviceGraph(kron14) as dg:
            res = at.tune_knobs(dg, algorithm, sources, space=space, repeats=1)
            assert all(t["exact"] for t in res["trials"])
            assert res["tuned_ms"] <= res["default_ms"]
            now = hc.tuning()
            for k, v in res["knobs"].items():
                assert now[k] == v
    finally:
        at.apply_knobs(before)


@pytest.mark.parametrize("algorithm", ["sssp", "bfs"])
def test_profiled_schedule_runs_exact(hc, gpu, oracle, at, kron14, algorithm):
    g = kron14
    sched, info = at.tune_schedule(g, algorithm, num_segments=6)
    prof = info["profile"]
    assert len(prof["segments"]) == 6
    assert all(s["cpu_ms"] > 0 and s["gpu"] for s in prof["segments"])
    src = int(g.pick_sources(1)[0])
    with hc.Hetero(g, sched) as h:
        if algorithm == "sssp":
            r = h.sssp(src)
            assert np.array_equal(r.values, oracle.sssp_dijkstra(g.index, g.neighbors,
                                                                  g.weights, src))
        else:
            r = h.bfs(src)
            assert np.array_equal(r.values, oracle.bfs_depth(g.index, g.neighbors, src))
    # A schedule that forces half the segments onto the host still runs.
    blocks = [{"device": "Host CPU", "start": 0, "end": 2,
               "kernel": at.CPU_KERNEL[algorithm]},
              {"device": "AMD Instinct MI355X", "start": 3, "end": 5,
               "kernel": at.GPU_KERNELS[algorithm][-1]}]
    with hc.Hetero(g, hc.Schedule.create(blocks, 6)) as h:
        r2 = h.sssp(src) if algorithm == "sssp" else h.bfs(src)
    assert np.array_equal(r.values, r2.values)


def test_tree_microbenchmark(hc, gpu, at, tmp_path):
    """The tree microbenchmark profiles 2^0..2^d segments per layer on the
    host and the GPU kernels and writes complete tree YAML files."""
    import yaml
    g = hc.HostGraph.generate("kron", 12, 16, weighted=True)
    tree = at.profile_tree(g, "sssp", 2, shapes=[(0, 0), (64, 256)], repeats=1)
    assert [len(l["segments"]) for l in tree["layers"]] == [1, 2, 4]
    for layer in tree["layers"]:
        assert all(s["cpu_ms"] > 0 and s["gpu"] for s in layer["segments"])
    files = at.write_yaml_tree(tree, g, str(tmp_path))
    assert len(files) >= 2
    for f in files:
        doc = yaml.safe_load(open(f))
        assert [r["depth"] for r in doc["results"]] == [0, 1, 2]
