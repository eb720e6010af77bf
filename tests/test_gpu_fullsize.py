"""Full-size GPU parity: every configuration of the workload at its real size.

Integer results are proven exact with the oracle's O(m) shortest-path
certificate (or_certify_paths, pinned in tests/test_certificate.py): a
distance or depth vector passes iff it equals the unique fixpoint that the
reference's CPU epoch loops (src/kernels/cpu/sssp.cuh:33-65,
src/kernels/cpu/bfs.cuh:48-114) and their serial oracles
(tests/sssp/sssp_verify.cu:122-146, tests/bfs/bfs_verify.cu:211-233) reach.
A few vectors per graph are also compared entry by entry with the oracle's
Dijkstra / FIFO BFS.  PageRank is compared with the FP32 reference iteration
and the FP64 oracle at the same iteration count.

  C2  Kronecker 23 and uniform 23 (ef 16, weights U[1,255]): BFS and SSSP from
      the first 64 SourcePicker sources, all certified; parents verified.
  C3  PageRank on Kronecker 24: L1 <= 1e-6 against the FP32 reference.
  C4a 64-source batched SSSP on Kronecker 22, and on uniform 23 (push epochs
      and buckets): all 64 columns certified.
  C4b 512 Kronecker graphs, scale U{10..16}, ef U{8..32}: every depth vector
      equal to the oracle's, every parent array valid.
  C5  serving over resident Kronecker 24 + uniform 23 and uploaded Kronecker
      12-18 graphs, Poisson arrivals: every result certified.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

SEED = 27491095      # graph seed of the benchmark
SOURCES = 64         # SourcePicker sources per graph (C2)
PR_L1_TOL = 1e-6     # PageRank: L1 over all vertices vs the FP32 reference


def _certify_columns(oracle, g, srcs, cols, weighted, unreached):
    """Certifies result vectors in groups of 64 columns; returns the
    per-column violation counts."""
    bad = []
    for i in range(0, len(cols), 64):
        M = np.ascontiguousarray(np.stack(cols[i:i + 64], axis=1))
        bad.extend(oracle.certify_paths(g.index, g.neighbors, g.weights if weighted else None,
                                        np.asarray(srcs[i:i + 64], np.int32), M, unreached))
        del M
    return np.array(bad)


@pytest.fixture(scope="module")
def c2(hc):
    out = {}
    for kind in ("kron", "uniform"):
        g = hc.HostGraph.generate(kind, 23, 16, weighted=True, seed=SEED)
        out[kind] = (g, hc.DeviceGraph(g))
    yield out
    for g, dg in out.values():
        dg.free()


@pytest.mark.parametrize("kind", ["kron", "uniform"])
def test_c2_sssp_64_sources(hc, gpu, oracle, c2, kind):
    g, dg = c2[kind]
    srcs = [int(s) for s in g.pick_sources(SOURCES)]
    buf = hc.DeviceArray(g.num_nodes, np.int32)
    cols = []
    for s in srcs:
        cols.append(hc.sssp(dg, s, out=buf).values)
    buf.free()
    bad = _certify_columns(oracle, g, srcs, cols, True, oracle.INF)
    assert bad.sum() == 0, (kind, np.flatnonzero(bad)[:8], bad[bad > 0][:8])
    for q in (0, SOURCES - 1):
        ref = oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, srcs[q])
        assert np.array_equal(cols[q], ref), (kind, srcs[q])


@pytest.mark.parametrize("kind", ["kron", "uniform"])
def test_c2_bfs_64_sources(hc, gpu, oracle, c2, kind):
    g, dg = c2[kind]
    srcs = [int(s) for s in g.pick_sources(SOURCES)]
    depth = hc.DeviceArray(g.num_nodes, np.int32)
    cols = []
    for s in srcs:
        r = hc.bfs(dg, s, depth_out=depth)
        cols.append(r.values)
        assert oracle.verify_parents(r.values, r.parents, s) == 0, (kind, s)
    depth.free()
    bad = _certify_columns(oracle, g, srcs, cols, False, -1)
    assert bad.sum() == 0, (kind, np.flatnonzero(bad)[:8])
    for q in (0, SOURCES - 1):
        assert np.array_equal(cols[q], oracle.bfs_depth(g.index, g.neighbors, srcs[q]))


def test_c3_pagerank_kron24(hc, gpu, oracle):
    g = hc.HostGraph.generate("kron", 24, 16, weighted=False, seed=SEED)
    with hc.DeviceGraph(g) as dg:
        p = hc.pagerank(dg)  # damping 0.85, L1 tolerance 1e-4, <= 20 iterations
    it = p.stats["epochs"]
    assert 1 <= it <= 20 and p.stats["residual"] < 1e-4
    got = p.values.astype(np.float64)
    ref32 = oracle.pr_jacobi_f32(g.index, g.neighbors, 0.85, it).astype(np.float64)
    ref64 = oracle.pr_jacobi_f64(g.index, g.neighbors, 0.85, it)
    l1_32, l1_64 = np.abs(got - ref32).sum(), np.abs(got - ref64).sum()
    print(f"kron24 pagerank iters={it} vs fp32 L1 {l1_32:.3e} max "
          f"{np.abs(got - ref32).max():.3e}; vs fp64 L1 {l1_64:.3e} max "
          f"{np.abs(got - ref64).max():.3e}")
    assert l1_32 <= PR_L1_TOL, l1_32
    assert l1_64 <= PR_L1_TOL, l1_64


def test_c4a_batched_sssp_kron22(hc, gpu, oracle):
    g = hc.HostGraph.generate("kron", 22, 16, weighted=True, seed=SEED)
    srcs = g.pick_sources(64)
    with hc.DeviceGraph(g) as dg:
        buf = hc.DeviceArray(g.num_nodes * hc.MULTI_WIDTH, np.int32)
        hc.sssp_multi(dg, srcs, out=buf, fetch=False)
        M = buf.get().reshape(g.num_nodes, hc.MULTI_WIDTH)
        buf.free()
        single = hc.sssp(dg, int(srcs[5])).values
    bad = oracle.certify_paths(g.index, g.neighbors, g.weights, srcs, M)
    assert bad.sum() == 0, np.flatnonzero(bad)
    assert np.array_equal(M[:, 5], single)
    assert np.array_equal(M[:, 0], oracle.sssp_dijkstra(g.index, g.neighbors, g.weights,
                                                        int(srcs[0])))


def test_c4a_batched_sssp_uniform23_buckets(hc, gpu, oracle, c2):
    """The batched sweep on the serving's flat graph (uniform 23: push epochs
    and delta-stepping buckets by default): every column certified, one
    equal to the single-source run (near-far buckets) and one to Dijkstra."""
    g, dg = c2["uniform"]
    srcs = g.pick_sources(64, seed=99)
    buf = hc.DeviceArray(g.num_nodes * hc.MULTI_WIDTH, np.int32)
    r = hc.sssp_multi(dg, srcs, out=buf, fetch=False)
    M = buf.get().reshape(g.num_nodes, hc.MULTI_WIDTH)
    buf.free()
    st = r.stats
    assert st["push_epochs"] > 0 and st["epochs"] > st["push_epochs"] + st["pull_epochs"], st
    bad = oracle.certify_paths(g.index, g.neighbors, g.weights, srcs, M)
    assert bad.sum() == 0, np.flatnonzero(bad)
    single = hc.sssp(dg, int(srcs[7]))
    assert single.stats["pull_epochs"] == 0  # near-far: push and promotion epochs only
    assert np.array_equal(M[:, 7], single.values)
    assert np.array_equal(M[:, 0], oracle.sssp_dijkstra(g.index, g.neighbors, g.weights,
                                                        int(srcs[0])))


def test_c4b_many_graph_bfs_512(hc, gpu, oracle):
    rng = np.random.RandomState(4242)
    graphs = [hc.HostGraph.generate("kron", int(rng.randint(10, 17)), int(rng.randint(8, 33)),
                                    weighted=False, seed=4242 + 1 + i) for i in range(512)]
    srcs = [int(g.pick_sources(1)[0]) for g in graphs]
    with hc.GraphBatch(graphs) as b:
        depths, parents, st = b.bfs(srcs)
    assert st["push_epochs"] + st["pull_epochs"] > 0
    for i, g in enumerate(graphs):
        want = oracle.bfs_depth(g.index, g.neighbors, srcs[i])
        assert np.array_equal(depths[i], want), i
        assert oracle.verify_parents(want, parents[i], srcs[i]) == 0, i


def test_c5_serving_fullsize(hc, gpu, oracle, c2):
    """Poisson arrivals (1000 q/s offered) of 70% BFS / 30% SSSP queries, 20%
    on small graphs uploaded on demand; every result is certified (the exact
    answer, hence the same as the query run alone), a sample is also
    compared with the query run alone on the GPU."""
    big = hc.HostGraph.generate("kron", 24, 16, weighted=True, seed=SEED)
    unif = c2["uniform"][0]
    rng = np.random.RandomState(777)
    dyn = [hc.HostGraph.generate("kron", int(s), 16, weighted=True, seed=SEED + int(s))
           for s in rng.randint(12, 19, size=8)]
    hosts = [big, unif] + dyn
    srv = hc.Server(slots=4, max_batch=64)
    gids = [srv.add_graph(big), srv.add_graph(unif)] + [srv.add_graph(g, resident=False)
                                                        for g in dyn]
    queries = 480
    qs = []
    for _ in range(queries):
        u = rng.rand()
        k = (2 + int(rng.randint(len(dyn)))) if u < 0.2 else (0 if u < 0.6 else 1)
        kind = hc.QUERY_BFS if rng.rand() < 0.7 else hc.QUERY_SSSP
        src = int(hosts[k].pick_sources(1, seed=int(rng.randint(1 << 30)))[0])
        qs.append((k, kind, src))
    arrive = np.cumsum(rng.exponential(1.0 / 1000, size=queries))
    t0 = time.perf_counter()
    tickets = []
    for i, (k, kind, src) in enumerate(qs):
        delay = t0 + arrive[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(gids[k], kind, src))
    groups = {}
    checked = 0
    alone = {0: [], 1: []}

    def flush(key):
        nonlocal checked
        k, kind = key
        srcs, cols = groups.pop(key)
        bad = _certify_columns(oracle, hosts[k], srcs, cols, kind == hc.QUERY_SSSP,
                               -1 if kind == hc.QUERY_BFS else oracle.INF)
        assert bad.sum() == 0, (k, kind, [srcs[j] for j in np.flatnonzero(bad)])
        checked += len(cols)

    batched = 0
    for i, t in enumerate(tickets):
        info = srv.wait(t)
        assert info["status"] == 0, info
        batched += info["batch_size"] > 1
        k, kind, src = qs[i]
        got = srv.fetch(t)
        srv.release(t)
        if k < 2 and len(alone[k]) < 3:
            alone[k].append((kind, src, got))
        key = (k, kind)
        groups.setdefault(key, ([], []))
        groups[key][0].append(src)
        groups[key][1].append(got)
        if len(groups[key][1]) == 64:
            flush(key)
    for key in list(groups):
        flush(key)
    st = srv.stats()
    srv.close()
    assert checked == queries and st["failed"] == 0 and st["completed"] == queries
    for k, items in alone.items():
        with hc.DeviceGraph(hosts[k]) as dg:
            for kind, src, got in items:
                ref = (hc.bfs(dg, src, want_parents=False).values if kind == hc.QUERY_BFS
                       else hc.sssp(dg, src).values)
                assert np.array_equal(got, ref), (k, kind, src)
    print(f"serving: {queries} queries certified, {batched} answered by batched sweeps, "
          f"uploads {st['uploads']} evictions {st['evictions']} spills {st['spills']}")
