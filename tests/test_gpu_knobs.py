"""GPU tests of the launch-shape knobs and of per-graph knobs: every
workgroup size of the batched sweep and of the many-graph BFS, and every
many-graph threshold, gives the oracle's results; knobs set on one graph
apply to it and its shares only, also from concurrent threads and through
the server."""
from __future__ import annotations

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kron13(hc):
    return hc.HostGraph.generate("kron", 13, 16, weighted=True, seed=31)


@pytest.fixture()
def restore(hc):
    yield
    hc.reset_tuning()


@pytest.mark.parametrize("block", [128, 256, 512, 1024])
def test_multi_block_sizes(hc, gpu, oracle, kron13, restore, block):
    g = kron13
    sources = g.pick_sources(64, seed=3)
    hc.set_tuning("multi.block", block)
    for per_cu in (1, 16):
        hc.set_tuning("multi.grid_per_cu", per_cu)
        with hc.DeviceGraph(g) as dg:
            r = hc.sssp_multi(dg, sources)
            r2 = hc.sssp_multi(dg, sources)  # relaunched epoch graph
        for q in (0, 21, 63):
            ref = oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, int(sources[q]))
            assert np.array_equal(r.values[q], ref), (block, per_cu, q)
        assert np.array_equal(r.values, r2.values)


def test_multi_block_change_rebuilds_epoch_graph(hc, gpu, kron13, restore):
    """The epoch graph bakes in the workgroup size: a new size on the same
    handle must not relaunch the old graph."""
    sources = kron13.pick_sources(64, seed=5)
    with hc.DeviceGraph(kron13) as dg:
        a = hc.sssp_multi(dg, sources)
        for block in (512, 128, 256):
            hc.set_tuning("multi.block", block)
            b = hc.sssp_multi(dg, sources)
            assert np.array_equal(a.values, b.values), block
            assert a.stats["epochs"] == b.stats["epochs"]


def _many_graphs(hc, count, seed):
    rng = np.random.RandomState(seed)
    return [hc.HostGraph.generate("kron" if i % 4 else "uniform", int(rng.randint(6, 14)),
                                  int(rng.randint(2, 33)), weighted=False, seed=500 + i)
            for i in range(count)]


def _check_many(oracle, graphs, sources, depths, parents):
    for i, (g, s) in enumerate(zip(graphs, sources)):
        ref = oracle.bfs_depth(g.index, g.neighbors, int(s))
        assert np.array_equal(depths[i], ref), i
        assert oracle.verify_parents(ref, parents[i], int(s)) == 0, i


@pytest.fixture(scope="module")
def many_set(hc):
    graphs = _many_graphs(hc, 40, 99)
    # One graph above 2^15 vertices: with many.lds_nodes below it, it takes
    # the global bitmaps while the others stay in LDS.
    graphs.append(hc.HostGraph.generate("kron", 16, 8, weighted=False, seed=7))
    sources = [int(g.pick_sources(1, seed=i)[0]) for i, g in enumerate(graphs)]
    return graphs, sources


@pytest.mark.parametrize("knob,values", [
    ("many.block", [256, 512, 1024]),
    ("many.per_cu", [0, 1, 4]),
    ("many.light_deg", [0, 8, 1 << 20]),
    ("many.heavy_cap", [0, 1, 1024]),
    ("many.lds_nodes", [0, 1 << 10, 1 << 15, 1 << 17]),
])
def test_many_knobs_exact(hc, gpu, oracle, many_set, restore, knob, values):
    graphs, sources = many_set
    with hc.GraphBatch(graphs) as b:
        base_d, _, base_st = b.bfs(sources)
        for v in values:
            hc.set_tuning(knob, v)
            d, p, st = b.bfs(sources)
            _check_many(oracle, graphs, sources, d, p)
            assert (st["push_epochs"], st["pull_epochs"]) == \
                (base_st["push_epochs"], base_st["pull_epochs"]), (knob, v)
            # The allocation-free entry point under the same knobs.
            d2, _, _ = b.bfs(sources, want_parents=False, noalloc=True)
            assert all(np.array_equal(x, y) for x, y in zip(d, d2)), (knob, v)
        hc.reset_tuning()


def test_many_block_1024_lds_off_all_global(hc, gpu, oracle, many_set, restore):
    graphs, sources = many_set
    hc.set_tuning("many.block", 1024)
    hc.set_tuning("many.lds_nodes", 0)
    hc.set_tuning("many.light_deg", 4)
    hc.set_tuning("many.heavy_cap", 16)
    with hc.GraphBatch(graphs) as b:
        d, p, _ = b.bfs(sources)
    _check_many(oracle, graphs, sources, d, p)


def test_knob_validation(hc, gpu, kron13, restore):
    for key, bad in (("multi.block", 192), ("many.block", 128), ("many.heavy_cap", 1025),
                     ("many.lds_nodes", (1 << 17) + 1), ("many.per_cu", -1)):
        with pytest.raises(hc.HcError):
            hc.set_tuning(key, bad)
    with hc.DeviceGraph(kron13) as dg:
        with pytest.raises(hc.HcError):
            dg.set_tuning("multi.block", 100)
        with pytest.raises(hc.HcError):
            dg.set_tuning("no.such.knob", 1)
        # A rejected value is not recorded.
        assert dg.get_tuning("multi.block") == hc.get_tuning("multi.block")


def test_graph_knobs_apply_to_graph_and_shares_only(hc, gpu, oracle, kron13, restore):
    g = kron13
    src = int(g.pick_sources(1)[0])
    ref = oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, src)
    with hc.DeviceGraph(g) as a, hc.DeviceGraph(g) as b:
        base = hc.sssp(b, src)
        # Pull whenever the frontier holds more than m/64 edges: more pull
        # epochs than the default on this graph.
        a.set_tuning("sssp.alpha", 64)
        assert a.get_tuning("sssp.alpha") == 64
        assert b.get_tuning("sssp.alpha") == hc.get_tuning("sssp.alpha")
        ra = hc.sssp(a, src)
        rb = hc.sssp(b, src)
        assert np.array_equal(ra.values, ref) and np.array_equal(rb.values, ref)
        assert ra.stats["pull_epochs"] != rb.stats["pull_epochs"]
        assert rb.stats["pull_epochs"] == base.stats["pull_epochs"]
        sh = a.share()  # inherits the graph's knobs
        rs = hc.sssp(sh, src)
        assert rs.stats["pull_epochs"] == ra.stats["pull_epochs"]
        assert sh.get_tuning("sssp.alpha") == 64
        # Process-wide knobs still reach the graph where it sets none.
        hc.set_tuning("profile.kernel_events", 1)
        rk = hc.sssp(a, src)
        hc.set_tuning("profile.kernel_events", 0)
        assert rk.stats["kernel_launches"] > 0
        a.reset_tuning()
        assert hc.sssp(a, src).stats["pull_epochs"] == base.stats["pull_epochs"]
        assert hc.sssp(sh, src).stats["pull_epochs"] == base.stats["pull_epochs"]
        sh.free()


def test_graph_knobs_from_concurrent_threads(hc, gpu, oracle, kron13, restore):
    """Two threads on two graphs with different knobs: each run sees its own
    graph's (the knobs are per call, not a process-wide switch)."""
    g = kron13
    srcs = [int(s) for s in g.pick_sources(6, seed=9)]
    refs = {s: oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, s) for s in srcs}
    with hc.DeviceGraph(g) as a, hc.DeviceGraph(g) as b:
        a.set_tuning("sssp.alpha", 64)
        b.set_tuning("sssp.alpha", 1)
        want = {}
        for name, dg in (("a", a), ("b", b)):
            want[name] = {s: hc.sssp(dg, s).stats["pull_epochs"] for s in srcs}
        errors = []

        def work(name, dg):
            try:
                for _ in range(3):
                    for s in srcs:
                        r = hc.sssp(dg, s)
                        st = r.stats
                        if not np.array_equal(r.values, refs[s]):
                            errors.append((name, s, "values"))
                        if st["pull_epochs"] != want[name][s]:
                            errors.append((name, s, st["pull_epochs"], want[name][s]))
            except Exception as e:  # reported below
                errors.append((name, repr(e)))

        ts = [threading.Thread(target=work, args=("a", a)),
              threading.Thread(target=work, args=("b", b))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errors, errors[:5]


def test_server_graph_knobs(hc, gpu, oracle, restore):
    """A knob recorded for a dynamic graph survives its eviction and
    re-upload; every result stays exact."""
    gs = [hc.HostGraph.generate("kron", 12, 8, weighted=True, seed=40 + i) for i in range(3)]
    srv = hc.Server(slots=2, max_batch=64)
    try:
        gids = [srv.add_graph(g, resident=False) for g in gs]
        for gid in gids:
            srv.set_graph_tuning(gid, "sssp.alpha", 64)
            srv.set_graph_tuning(gid, "multi.block", 512)
        with pytest.raises(hc.HcError):
            srv.set_graph_tuning(gids[0], "multi.block", 100)
        with pytest.raises(hc.HcError):
            srv.set_graph_tuning(99, "sssp.alpha", 4)
        tickets = []
        for rep in range(2):
            for gid, g in zip(gids, gs):
                for s in g.pick_sources(5, seed=rep):
                    tickets.append((gid, g, int(s), srv.submit(gid, hc.QUERY_SSSP, int(s))))
        for gid, g, s, t in tickets:
            info = srv.wait(t)
            assert info["status"] == 0, info
            got = srv.fetch(t)
            srv.release(t)
            assert np.array_equal(got, oracle.sssp_dijkstra(g.index, g.neighbors, g.weights,
                                                            s)), (gid, s)
    finally:
        srv.close()
