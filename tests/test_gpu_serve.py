"""GPU tests of the serving loop: continuous batching, admission by memory
footprint and dynamic graphs.  Every result must be bit-identical to the
same query run alone, and to the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import result_checksum

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs(hc):
    return {
        "kron": hc.HostGraph.generate("kron", 14, 16, weighted=True, seed=1),
        "unif": hc.HostGraph.generate("uniform", 13, 16, weighted=True, seed=2),
        "dyn": [hc.HostGraph.generate("kron", s, 8, weighted=True, seed=10 + s)
                for s in (10, 11, 12)],
    }


def _workload(graphs, ids, count, seed):
    rng = np.random.RandomState(seed)
    qs = []
    for _ in range(count):
        r = rng.rand()
        if r < 0.2:
            k = int(rng.randint(len(graphs["dyn"])))
            gid, g = ids["dyn"][k], graphs["dyn"][k]
        elif r < 0.6:
            gid, g = ids["kron"], graphs["kron"]
        else:
            gid, g = ids["unif"], graphs["unif"]
        kind = 0 if rng.rand() < 0.7 else 1
        src = int(g.pick_sources(1, seed=int(rng.randint(1 << 30)))[0])
        qs.append((gid, kind, src, g))
    return qs


def _server(hc, graphs, **kw):
    srv = hc.Server(**kw)
    ids = {"kron": srv.add_graph(graphs["kron"]), "unif": srv.add_graph(graphs["unif"]),
           "dyn": [srv.add_graph(g, resident=False) for g in graphs["dyn"]]}
    return srv, ids


def _answers(srv, tickets):
    """Fetches every result; its device checksum (hc_server_checksum, kept
    from the sweep or computed where the result lies, spilled or not) must
    match the restatement's over the fetched values."""
    out = []
    for t in tickets:
        info = srv.wait(t)
        assert info["status"] == 0, info
        got = srv.fetch(t)
        assert srv.checksum(t) == result_checksum(got), (t, info)
        out.append((got, info))
        srv.release(t)
    return out


def test_mix_equals_alone(hc, gpu, oracle, graphs):
    # Alone: one slot, no batching, each query waited for before the next.
    srv, ids = _server(hc, graphs, slots=1, max_batch=1)
    qs = _workload(graphs, ids, 80, 7)
    alone = []
    for gid, kind, src, _ in qs:
        t = srv.submit(gid, kind, src)
        alone.append(_answers(srv, [t])[0][0])
    srv.close()
    # Mix: every query submitted at once to a multi-slot server that
    # batches any two or more waiting queries of one graph and kind.
    srv, ids = _server(hc, graphs, slots=4, max_batch=64, min_batch_sssp=2, min_batch_bfs=2,
                       checksum=True)
    tickets = [srv.submit(gid, kind, src) for gid, kind, src, _ in qs]
    mixed = _answers(srv, tickets)
    st = srv.stats()
    srv.close()
    assert max(info["batch_size"] for _, info in mixed) > 1, "no query was batched"
    assert st["batched_queries"] > 0 and st["completed"] == len(qs)
    for i, ((gid, kind, src, g), a, (m, info)) in enumerate(zip(qs, alone, mixed)):
        assert np.array_equal(a, m), (i, kind, info["batch_size"])
        if i % 8 == 0:  # oracle on a sample
            ref = (oracle.bfs_depth(g.index, g.neighbors, src) if kind == 0
                   else oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, src))
            assert np.array_equal(m, ref), (i, kind)


def test_memory_budget_evicts_dynamic_graphs(hc, gpu, oracle, graphs):
    # Budget: room for one 64-wide result of the largest dynamic graph and
    # little else, so dynamic graphs are evicted and re-uploaded.  The
    # device memory the server really holds (free-memory delta, taken after
    # the queries with every uploaded graph and pooled buffer still held)
    # must stay within the budget.
    probe = hc.Server(slots=2)
    g_max = max(probe.graph_bytes(g, False) for g in graphs["dyn"])
    probe.close()
    n_max = max(g.num_nodes for g in graphs["dyn"])
    budget = n_max * 4 * 64 + g_max + 65536
    srv = hc.Server(slots=2, max_batch=64, memory_budget=budget, min_batch_sssp=2,
                    min_batch_bfs=2)
    # Baseline taken with the server's streams created: their queues are the
    # runtime's, not graph or result memory, and sit outside the budget.
    hc.synchronize()
    free0, _ = hc.mem_get_info()
    dyn = [srv.add_graph(g, resident=False) for g in graphs["dyn"]]
    rng = np.random.RandomState(3)
    qs = []
    for r in range(3):
        for k, g in enumerate(graphs["dyn"]):
            for s in g.pick_sources(4, seed=100 * r + k):
                qs.append((dyn[k], int(rng.randint(2)), int(s), g))
    tickets = [srv.submit(gid, kind, src) for gid, kind, src, _ in qs]
    res = _answers(srv, tickets)
    st = srv.stats()
    hc.synchronize()
    free1, _ = hc.mem_get_info()
    srv.close()
    assert st["evictions"] > 0 and st["uploads"] > len(graphs["dyn"])
    assert st["bytes_peak"] <= budget
    # the fetch staging buffer (one column) sits outside the budget: 8 MiB slack
    assert free0 - free1 <= budget + (8 << 20), (free0 - free1, budget)
    for (gid, kind, src, g), (m, _) in zip(qs, res):
        ref = (oracle.bfs_depth(g.index, g.neighbors, src) if kind == 0
               else oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, src))
        assert np.array_equal(m, ref)


def test_below_batch_threshold_runs_single_source(hc, gpu, oracle, graphs):
    """Fewer waiting queries of one graph and kind than the batching
    threshold run one at a time on the single-source loops: with the BFS
    threshold at 64, twenty BFS queries submitted at once all answer from
    runs of their own, exactly; SSSP queries (threshold 2) may share a
    sweep and stay exact."""
    srv = hc.Server(slots=2, max_batch=64, min_batch_sssp=2, min_batch_bfs=64)
    g = graphs["unif"]
    gid = srv.add_graph(g)
    srcs = [int(x) for x in g.pick_sources(20, seed=11)]
    tickets = [srv.submit(gid, 0, s) for s in srcs] + [srv.submit(gid, 1, s) for s in srcs[:8]]
    res = _answers(srv, tickets)
    srv.close()
    for i, (m, info) in enumerate(res):
        s = srcs[i % 20] if i < 20 else srcs[i - 20]
        if i < 20:
            assert info["batch_size"] == 1, info
            assert np.array_equal(m, oracle.bfs_depth(g.index, g.neighbors, s))
        else:
            assert np.array_equal(m, oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, s))


def test_server_rejects_bad_queries(hc, gpu, graphs):
    srv = hc.Server(slots=1)
    gid = srv.add_graph(graphs["kron"])
    with pytest.raises(hc.HcError):
        srv.submit(gid, 0, graphs["kron"].num_nodes)
    with pytest.raises(hc.HcError):
        srv.submit(gid + 5, 0, 0)
    with pytest.raises(hc.HcError):
        srv.wait(12345)
    srv.close()


def test_fast_sssp_arrivals_gather_batches(hc, gpu, oracle, graphs):
    """SSSP queries arriving fast enough to gather 2 x min_batch within
    max_wait_ms wait for a batch; BFS queries, whose fixed threshold is a
    full batch, run alone or as full batches; sparse arrivals run at once,
    one query per run.  Every answer exact.  (Fixed thresholds: no cost
    calibration.)"""
    import time
    g = graphs["kron"]
    srv = hc.Server(slots=4, max_batch=64, max_wait_ms=300, calibrate=-1)
    gid = srv.add_graph(g)
    rng = np.random.RandomState(5)
    fast = [(int(rng.rand() < 0.3), int(s)) for s in g.pick_sources(240, seed=9)]
    t0 = time.perf_counter()
    tickets = []
    for i, (kind, src) in enumerate(fast):       # ~2000 queries/s
        delay = t0 + i * 0.0005 - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(gid, kind, src))
    res = _answers(srv, tickets)
    sssp_batches = [info["batch_size"] for (kind, _), (_, info) in zip(fast, res) if kind == 1]
    bfs_batches = [info["batch_size"] for (kind, _), (_, info) in zip(fast, res) if kind == 0]
    assert np.mean(sssp_batches) >= 16, sssp_batches
    assert all(b == 1 or b == 64 for b in bfs_batches), bfs_batches
    for (kind, src), (m, info) in zip(fast, res):
        ref = (oracle.bfs_depth(g.index, g.neighbors, src) if kind == 0
               else oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, src))
        assert np.array_equal(m, ref), (kind, src)
    # Sparse arrivals (one SSSP query every 60 ms): no gathering.
    slow = [int(s) for s in g.pick_sources(6, seed=21)]
    tickets = []
    for src in slow:
        tickets.append(srv.submit(gid, 1, src))
        time.sleep(0.06)
    res = _answers(srv, tickets)
    assert all(info["batch_size"] == 1 for _, info in res)
    assert all(info["queue_ms"] < 200 for _, info in res), [info["queue_ms"] for _, info in res]
    for src, (m, _) in zip(slow, res):
        assert np.array_equal(m, oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, src))
    srv.close()


def test_resident_graph_costs_measured(hc, gpu, oracle):
    """Adding a resident graph times a single query and a 64-wide batch of
    each kind; calibrate=-1 skips it; queries stay exact either way."""
    g = hc.HostGraph.generate("kron", 13, 16, weighted=True, seed=12)
    u = hc.HostGraph.generate("kron", 12, 8, weighted=False, seed=13)
    with hc.Server(slots=2) as srv:
        a = srv.add_graph(g)
        b = srv.add_graph(u)
        d = srv.add_graph(hc.HostGraph.generate("kron", 11, 8, weighted=True, seed=14),
                          resident=False)
        ca, cb, cd = srv.graph_costs(a), srv.graph_costs(b), srv.graph_costs(d)
        for k in ("bfs", "sssp"):
            assert ca[k]["single_ms"] > 0 and ca[k]["batch_ms"] > 0
        assert cb["bfs"]["single_ms"] > 0 and cb["sssp"]["single_ms"] == 0  # unweighted
        assert all(v == 0 for c in cd.values() for v in c.values())  # dynamic: unmeasured
        tickets = [(gid, host, kind, int(s), srv.submit(gid, kind, int(s)))
                   for gid, host in ((a, g), (b, u))
                   for kind in ((hc.QUERY_BFS, hc.QUERY_SSSP) if host.weights is not None
                                else (hc.QUERY_BFS,))
                   for s in host.pick_sources(20, seed=kind)]
        for gid, host, kind, s, t in tickets:
            got = srv.fetch(t)
            srv.release(t)
            ref = (oracle.bfs_depth(host.index, host.neighbors, s) if kind == hc.QUERY_BFS
                   else oracle.sssp_dijkstra(host.index, host.neighbors, host.weights, s))
            assert np.array_equal(got, ref), (gid, kind, s)
    with hc.Server(slots=1, calibrate=-1) as srv:
        a = srv.add_graph(g)
        assert all(v == 0 for c in srv.graph_costs(a).values() for v in c.values())


def test_checksums_mix_vs_alone_without_fetch(hc, gpu, graphs):
    """The bench's check of every resident result: checksums kept after each
    sweep of a batching server equal those of the same queries run one at a
    time (single-source path), with no result copied to the host."""
    g = graphs["kron"]
    srcs = [int(x) for x in g.pick_sources(40, seed=31)]
    with hc.Server(slots=3, max_batch=64, min_batch_sssp=2, min_batch_bfs=2, max_wait_ms=-1,
                   checksum=True) as srv:
        gid = srv.add_graph(g)
        tickets = [(kind, s, srv.submit(gid, kind, s)) for s in srcs for kind in (0, 1)]
        mixed = {}
        for kind, s, t in tickets:
            info = srv.wait(t)
            mixed[(kind, s)] = (srv.checksum(t), info["batch_size"])
            srv.release(t)
        assert max(b for _, b in mixed.values()) > 1
        for (kind, s), (h, _) in mixed.items():
            t = srv.submit(gid, kind, s)
            assert srv.wait(t)["batch_size"] == 1
            assert srv.checksum(t) == h, (kind, s)
            srv.release(t)
