"""GPU parity of segment-scheduled co-execution: schedules that put some
segments on the GPU (range handles, one stream per block) and the rest on
the host cores, in the reference's .skd text.  SSSP must be bit-exact with
Dijkstra, BFS depths exact with valid parents and the oracle's push/pull
epoch counts, PageRank within 1e-6 L1 of the FP32 reference iteration.

Reference: src/kernels/heterogeneous/{sssp,pr,bfs}.cuh."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden
from test_schedule import REFERENCE_LAYOUT

pytestmark = pytest.mark.gpu

PR_L1_TOL = 1e-6


def _fixture_graph(hc, name):
    fx = load_golden(name)
    return fx, hc.HostGraph.from_csr(fx["index"], fx["neighbors"], fx["weights"])


@pytest.fixture(scope="module")
def kron14(hc):
    return hc.HostGraph.generate("kron", 14, 16, weighted=True)


def _mixed(hc, algo, gpu_kernel):
    cpu = {"sssp": "SSSP CPU one-to-one", "pr": "PR CPU one-to-one", "bfs": "BFS CPU pull"}[algo]
    return hc.Schedule.create([
        {"device": "AMD Instinct MI355X", "start": 0, "end": 2, "kernel": gpu_kernel},
        {"device": "Host CPU", "start": 3, "end": 3, "kernel": cpu},
        {"device": "AMD Instinct MI355X", "start": 4, "end": 6, "kernel": gpu_kernel,
         "blocks": 64, "threads": 256},
        {"device": "Host CPU", "start": 7, "end": 7, "kernel": cpu},
    ], 8)


@pytest.mark.parametrize("name", ["kron8.npz", "kron10.npz", "disconnected.npz",
                                  "path10.npz"])
def test_reference_layout_sssp_golden(hc, gpu, name):
    fx, g = _fixture_graph(hc, name)
    with hc.Hetero(g, hc.Schedule.parse(REFERENCE_LAYOUT)) as h:
        for i, s in enumerate(fx["sources"]):
            r = h.sssp(int(s))
            assert np.array_equal(r.values, fx["sssp"][i]), (name, s)


@pytest.mark.parametrize("kernel", ["SSSP GPU one-to-one", "SSSP GPU warp-red",
                                    "SSSP GPU block-red", "SSSP GPU bucketed"])
def test_mixed_sssp_kron14(hc, gpu, oracle, kron14, kernel):
    g = kron14
    with hc.Hetero(g, _mixed(hc, "sssp", kernel)) as h:
        for s in g.pick_sources(3):
            s = int(s)
            r = h.sssp(s)
            ref = oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, s)
            assert np.array_equal(r.values, ref), (kernel, s)
            assert r.stats["epochs"] >= 2


def test_gpu_only_schedule_sssp(hc, gpu, oracle, kron14):
    g = kron14
    sched = hc.Schedule.parse("AMD Instinct MI355X\n0 5\nSSSP GPU bucketed\n")
    with hc.Hetero(g, sched) as h:
        s = int(g.pick_sources(1)[0])
        r = h.sssp(s)
    assert np.array_equal(r.values, oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, s))


@pytest.mark.parametrize("kernel", ["BFS GPU one-to-one", "BFS GPU warp", "BFS GPU bucketed"])
def test_mixed_bfs_kron14(hc, gpu, oracle, kron14, kernel):
    g = kron14
    with hc.Hetero(g, _mixed(hc, "bfs", kernel)) as h:
        for s in g.pick_sources(3):
            s = int(s)
            r = h.bfs(s)
            depth = oracle.bfs_depth(g.index, g.neighbors, s)
            assert np.array_equal(r.values, depth), (kernel, s)
            assert oracle.verify_parents(depth, r.parents, s) == 0
            _, _, dirs = oracle.bfs_do(g.index, g.neighbors, s)
            assert r.stats["pull_epochs"] == int((dirs == 1).sum())
            assert r.stats["push_epochs"] == int((dirs == 0).sum())
            assert r.stats["pull_epochs"] >= 1


@pytest.mark.parametrize("kernel", ["PR GPU one-to-one", "PR GPU bucketed"])
def test_mixed_pagerank_kron14(hc, gpu, oracle, kron14, kernel):
    g = kron14
    with hc.Hetero(g, _mixed(hc, "pr", kernel)) as h:
        r = h.pagerank(tolerance=-1.0, max_iters=20)
    assert r.stats["epochs"] == 20
    ref = oracle.pr_jacobi_f32(g.index, g.neighbors, 0.85, 20)
    assert np.abs(r.values.astype(np.float64) - ref).sum() <= PR_L1_TOL


def test_mixed_pagerank_tolerance_stop(hc, gpu, kron14):
    with hc.DeviceGraph(kron14) as dg:
        full = hc.pagerank(dg)
    with hc.Hetero(kron14, _mixed(hc, "pr", "PR GPU bucketed")) as h:
        r = h.pagerank()
    assert r.stats["residual"] < 1e-4
    assert abs(r.stats["epochs"] - full.stats["epochs"]) <= 1
    assert np.abs(r.values.astype(np.float64) - full.values).sum() <= 1e-5
