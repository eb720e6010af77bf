"""GPU parity of batched queries: 64-source SSSP in one device-driven loop
must give, for every source, exactly the distances of Dijkstra (and of the
single-source GPU run)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def kron14(hc):
    return hc.HostGraph.generate("kron", 14, 16, weighted=True)


def _check_columns(oracle, g, sources, vals):
    for q, s in enumerate(sources):
        ref = oracle.sssp_dijkstra(g.index, g.neighbors, g.weights, int(s))
        assert np.array_equal(vals[q], ref), (q, int(s))


def test_multi_64_sources_kron14(hc, gpu, oracle, kron14):
    g = kron14
    sources = g.pick_sources(64)
    with hc.DeviceGraph(g) as dg:
        r = hc.sssp_multi(dg, sources)
        _check_columns(oracle, g, sources, r.values)
        assert r.stats["epochs"] >= 2
        # Same columns as the single-source runs.
        for q in (0, 17, 63):
            assert np.array_equal(hc.sssp(dg, int(sources[q])).values, r.values[q])


def test_multi_graph_and_plain_launches_agree(hc, gpu, kron14):
    sources = kron14.pick_sources(64, seed=7)
    with hc.DeviceGraph(kron14) as dg:
        a = hc.sssp_multi(dg, sources)
        hc.set_tuning("multi.graph", 0)
        try:
            b = hc.sssp_multi(dg, sources)
        finally:
            hc.set_tuning("multi.graph", 1)
        c = hc.sssp_multi(dg, sources)  # captured graph relaunched
    assert np.array_equal(a.values, b.values)
    assert np.array_equal(a.values, c.values)


@pytest.mark.parametrize("count", [1, 5, 33])
def test_multi_partial_batches_and_duplicates(hc, gpu, oracle, kron14, count):
    g = kron14
    sources = g.pick_sources(count, seed=11)
    if count > 2:
        sources[1] = sources[0]  # the same source in two columns
    with hc.DeviceGraph(g) as dg:
        out = hc.DeviceArray(g.num_nodes * hc.MULTI_WIDTH, np.int32)
        r = hc.sssp_multi(dg, sources, out=out)
        full = out.get().reshape(g.num_nodes, hc.MULTI_WIDTH)
    _check_columns(oracle, g, sources, r.values)
    assert np.all(full[:, count:] == hc.INF_DIST)


@pytest.mark.parametrize("name", ["kron10.npz", "disconnected.npz", "multi.npz",
                                  "zeroweight.npz", "path10.npz", "star9.npz"])
def test_multi_golden(hc, gpu, name):
    fx = load_golden(name)
    g = hc.HostGraph.from_csr(fx["index"], fx["neighbors"], fx["weights"])
    with hc.DeviceGraph(g) as dg:
        r = hc.sssp_multi(dg, fx["sources"])
    for q in range(len(fx["sources"])):
        assert np.array_equal(r.values[q], fx["sssp"][q]), (name, q)


@pytest.mark.parametrize("delta,push_div", [(1, 0), (5, 16), (32, 1), (1 << 20, 1 << 20),
                                            (0, 1), (0, 0)])
@pytest.mark.parametrize("kind,scale", [("uniform", 13), ("kron", 13)])
def test_multi_bucketed_and_push_epochs(hc, gpu, oracle, kind, scale, delta, push_div):
    """Improvements at or above a rising threshold wait (multi.delta), and
    epochs whose flagged vertices hold few edges push from them
    (multi.push_div): every column still equals Dijkstra's and the plain
    sweep's; small widths take promotion launches, push_div 0 none of
    push epochs, push_div 1 pushes from the sources at least."""
    g = hc.HostGraph.generate(kind, scale, 16, weighted=True, seed=delta % 89)
    sources = g.pick_sources(64, seed=delta % 13)
    sources[5] = sources[4]
    try:
        with hc.DeviceGraph(g) as dg:
            hc.set_tuning("multi.delta", 0)
            hc.set_tuning("multi.push_div", 0)
            plain = hc.sssp_multi(dg, sources)
            hc.set_tuning("multi.delta", delta)
            hc.set_tuning("multi.push_div", push_div)
            r = hc.sssp_multi(dg, sources)
            again = hc.sssp_multi(dg, sources)
    finally:
        hc.reset_tuning()
    assert np.array_equal(r.values, plain.values)
    assert np.array_equal(again.values, r.values)
    _check_columns(oracle, g, sources, r.values)
    st, ps = r.stats, plain.stats
    assert ps["epochs"] == ps["pull_epochs"] and ps["push_epochs"] == 0
    promotions = st["epochs"] - st["pull_epochs"] - st["push_epochs"]
    assert promotions >= 0
    if 0 < delta <= 5:
        assert promotions > 0, st
    if delta == 0:
        assert promotions == 0, st
    if push_div == 0:
        assert st["push_epochs"] == 0, st
    if push_div == 1 and np.diff(g.index)[sources].max() <= 1024:
        assert st["push_epochs"] > 0, st


@pytest.mark.parametrize("name", ["kron10.npz", "disconnected.npz", "zeroweight.npz",
                                  "path10.npz", "star9.npz"])
def test_multi_bucketed_golden(hc, gpu, name):
    """Tiny graphs with a bucket width of 1 (more launches than vertices)
    and every epoch a push."""
    fx = load_golden(name)
    g = hc.HostGraph.from_csr(fx["index"], fx["neighbors"], fx["weights"])
    try:
        hc.set_tuning("multi.delta", 1)
        hc.set_tuning("multi.push_div", 1)
        with hc.DeviceGraph(g) as dg:
            r = hc.sssp_multi(dg, fx["sources"])
    finally:
        hc.reset_tuning()
    for q in range(len(fx["sources"])):
        assert np.array_equal(r.values[q], fx["sssp"][q]), (name, q)


def test_multi_unsorted_ids_and_wide_offsets(hc, gpu, oracle, kron14):
    """Vertex ids not in degree order (plan lists instead of ranges) and
    8-byte offsets."""
    g = kron14
    n = g.num_nodes
    perm = np.random.default_rng(5).permutation(n).astype(np.int32)
    deg = np.diff(g.index)
    src = perm[np.repeat(np.arange(n), deg)]
    dst = perm[g.neighbors]
    h = hc.HostGraph.from_edges(src, dst, n, weights=g.weights, symmetrize=False,
                                relabel=False)
    sources = perm[g.pick_sources(64, seed=3)]
    os.environ["HC_FORCE_OFF64"] = "1"
    try:
        dg = hc.DeviceGraph(h)
    finally:
        del os.environ["HC_FORCE_OFF64"]
    with dg:
        info = dg.info()
        assert info["offset_bytes"] == 8 and not info["degree_sorted"]
        r = hc.sssp_multi(dg, sources)
    _check_columns(oracle, h, sources, r.values)


def test_multi_rejects_bad_input(hc, gpu, kron14):
    with hc.DeviceGraph(kron14) as dg:
        with pytest.raises(hc.HcError):
            hc.sssp_multi(dg, np.arange(65))
        with pytest.raises(hc.HcError):
            hc.sssp_multi(dg, [kron14.num_nodes])


def _graph_set(hc, count, seed, scales=(6, 12), efs=(2, 32)):
    rng = np.random.RandomState(seed)
    graphs = []
    for i in range(count):
        kind = "kron" if i % 3 else "uniform"
        graphs.append(hc.HostGraph.generate(kind, int(rng.randint(scales[0], scales[1] + 1)),
                                            int(rng.randint(efs[0], efs[1] + 1)),
                                            weighted=False, seed=1000 + i))
    return graphs


def _check_many(oracle, graphs, sources, depths, parents, stats):
    pushes = pulls = 0
    for i, (g, s) in enumerate(zip(graphs, sources)):
        ref = oracle.bfs_depth(g.index, g.neighbors, int(s))
        assert np.array_equal(depths[i], ref), i
        assert oracle.verify_parents(ref, parents[i], int(s)) == 0, i
        _, _, dirs = oracle.bfs_do(g.index, g.neighbors, int(s))
        pushes += int((dirs == 0).sum())
        pulls += int((dirs == 1).sum())
    assert stats["push_epochs"] == pushes
    assert stats["pull_epochs"] == pulls


def test_bfs_many_mixed_sizes(hc, gpu, oracle):
    graphs = _graph_set(hc, 48, 2024)
    sources = [int(g.pick_sources(1, seed=i)[0]) for i, g in enumerate(graphs)]
    with hc.GraphBatch(graphs) as b:
        assert b.total_edges == sum(g.num_edges for g in graphs)
        depths, parents, st = b.bfs(sources)
        _check_many(oracle, graphs, sources, depths, parents, st)
        # A second call with other sources reuses the resident batch.
        sources2 = [int(g.pick_sources(2, seed=i)[1]) for i, g in enumerate(graphs)]
        depths2, parents2, st2 = b.bfs(sources2)
        _check_many(oracle, graphs, sources2, depths2, parents2, st2)


def test_bfs_many_matches_single_graph_runs(hc, gpu):
    graphs = _graph_set(hc, 6, 77, scales=(10, 13), efs=(8, 32))
    sources = [int(g.pick_sources(1)[0]) for g in graphs]
    with hc.GraphBatch(graphs) as b:
        depths, _, st = b.bfs(sources, want_parents=False)
    for g, s, d in zip(graphs, sources, depths):
        with hc.DeviceGraph(g) as dg:
            assert np.array_equal(hc.bfs(dg, s).values, d)


def test_bfs_many_large_graph_global_bitmaps(hc, gpu, oracle):
    """A graph above the LDS bitmap limit (2^17 vertices) next to small ones."""
    big = hc.HostGraph.generate("kron", 18, 4, weighted=False, seed=5)
    graphs = [big] + _graph_set(hc, 5, 9)
    sources = [int(g.pick_sources(1)[0]) for g in graphs]
    with hc.GraphBatch(graphs) as b:
        depths, parents, st = b.bfs(sources)
    _check_many(oracle, graphs, sources, depths, parents, st)


def test_bfs_many_golden_and_isolated_source(hc, gpu):
    names = ["path10.npz", "star9.npz", "cycle12.npz", "complete6.npz", "disconnected.npz"]
    fxs = [load_golden(n) for n in names]
    graphs = [hc.HostGraph.from_csr(fx["index"], fx["neighbors"]) for fx in fxs]
    sources = [int(fx["sources"][0]) for fx in fxs]
    # An isolated vertex as source: only itself at depth 0.
    iso = hc.HostGraph.from_edges([0, 1], [1, 2], 5)
    graphs.append(iso)
    sources.append(4)
    with hc.GraphBatch(graphs) as b:
        depths, parents, _ = b.bfs(sources)
    for i, fx in enumerate(fxs):
        assert np.array_equal(depths[i], fx["depth"][0]), names[i]
    assert depths[-1].tolist() == [-1, -1, -1, -1, 0]
    assert parents[-1].tolist() == [-1, -1, -1, -1, 4]


def test_bfs_many_rejects_bad_sources(hc, gpu):
    graphs = _graph_set(hc, 2, 1)
    with hc.GraphBatch(graphs) as b:
        with pytest.raises(hc.HcError):
            b.bfs([0, graphs[1].num_nodes])


@pytest.mark.parametrize("weighted", [False, True])
def test_bfs_multi_depths(hc, gpu, oracle, weighted):
    g = hc.HostGraph.generate("kron", 14, 16, weighted=weighted)
    sources = g.pick_sources(64, seed=21)
    with hc.DeviceGraph(g) as dg:
        r = hc.bfs_multi(dg, sources)
        for q in (0, 1, 40, 63):
            assert np.array_equal(hc.bfs(dg, int(sources[q])).values, r.values[q])
    for q, s in enumerate(sources):
        assert np.array_equal(r.values[q], oracle.bfs_depth(g.index, g.neighbors, int(s))), q


def test_bfs_multi_disconnected_golden(hc, gpu):
    fx = load_golden("disconnected.npz")
    g = hc.HostGraph.from_csr(fx["index"], fx["neighbors"])
    with hc.DeviceGraph(g) as dg:
        r = hc.bfs_multi(dg, fx["sources"])
    for q in range(len(fx["sources"])):
        assert np.array_equal(r.values[q], fx["depth"][q])


# Bit-parallel batched BFS (msbfs.on = 1, the default) against the batched
# sweep with unit weights (msbfs.on = 0) and the oracle, with each level
# direction forced in turn (msbfs.alpha 1: push while the frontier has fewer
# edges than the graph; 2^20: pull from the first level on).

@pytest.fixture()
def reset_knobs(hc):
    yield
    hc.reset_tuning()


def _bfs_multi_both(hc, dg, sources, max_epochs=0):
    hc.set_tuning("msbfs.on", 0)
    try:
        old = hc.bfs_multi(dg, sources, max_epochs=max_epochs)
    finally:
        hc.set_tuning("msbfs.on", 1)
    return hc.bfs_multi(dg, sources, max_epochs=max_epochs), old


@pytest.mark.parametrize("alpha", [1, 8, 1 << 20])
@pytest.mark.parametrize("kind,scale", [("kron", 14), ("uniform", 13)])
def test_bfs_multi_bitwise_levels(hc, gpu, oracle, reset_knobs, alpha, kind, scale):
    g = hc.HostGraph.generate(kind, scale, 16, weighted=False, seed=3)
    sources = g.pick_sources(64, seed=alpha % 97)
    hc.set_tuning("msbfs.alpha", alpha)
    with hc.DeviceGraph(g) as dg:
        new, old = _bfs_multi_both(hc, dg, sources)
    assert np.array_equal(new.values, old.values)
    for q in (0, 9, 63):
        assert np.array_equal(new.values[q], oracle.bfs_depth(g.index, g.neighbors,
                                                               int(sources[q]))), q
    st = new.stats
    assert st["epochs"] == st["push_epochs"] + st["pull_epochs"]
    # The last level finds nothing: levels = largest depth + 1.
    assert st["epochs"] == int(new.values.max()) + 1
    if alpha == 1:
        # Pull only when a level's frontier holds every edge (the 64
        # frontiers can cover the whole graph at once).
        assert st["pull_epochs"] <= 1
    if alpha == 1 << 20:
        assert st["push_epochs"] == 0


@pytest.mark.parametrize("count", [1, 2, 7, 33])
def test_bfs_multi_bitwise_partial_and_duplicates(hc, gpu, oracle, reset_knobs, count):
    g = hc.HostGraph.generate("kron", 13, 8, weighted=False, seed=17)
    sources = g.pick_sources(count, seed=count)
    if count > 2:
        sources[2] = sources[0]  # one vertex in two columns
    with hc.DeviceGraph(g) as dg:
        d = hc.DeviceArray(g.num_nodes * hc.MULTI_WIDTH, np.int32)
        r = hc.bfs_multi(dg, sources, out=d)
        raw = d.get().reshape(g.num_nodes, hc.MULTI_WIDTH)
        d.free()
    for q, s in enumerate(sources):
        assert np.array_equal(r.values[q], oracle.bfs_depth(g.index, g.neighbors, int(s))), q
    # Columns past num_sources stay unreached.
    assert np.all(raw[:, count:] == hc.INF_DIST)


def test_bfs_multi_bitwise_isolated_sources_and_limit(hc, gpu, oracle, reset_knobs):
    g = hc.HostGraph.generate("kron", 12, 4, weighted=False, seed=5)
    deg = np.diff(g.index)
    iso = np.flatnonzero(deg == 0)[:3]
    live = g.pick_sources(5)
    sources = np.concatenate([iso, live]).astype(np.int32)
    with hc.DeviceGraph(g) as dg:
        full = hc.bfs_multi(dg, sources)
        lim, old_lim = _bfs_multi_both(hc, dg, sources, max_epochs=2)
    for q, s in enumerate(sources):
        ref = oracle.bfs_depth(g.index, g.neighbors, int(s))
        assert np.array_equal(full.values[q], ref), q
        # At most two levels: depths up to 2 exact, farther vertices unreached.
        want = np.where(ref <= 2, ref, -1)
        assert np.array_equal(lim.values[q], want), q
    assert lim.stats["epochs"] == 2


def test_bfs_multi_bitwise_wide_offsets_and_unsorted(hc, gpu, oracle, reset_knobs,
                                                     monkeypatch):
    g = hc.HostGraph.generate("kron", 12, 16, weighted=False, seed=8)
    rng = np.random.default_rng(2)
    perm = rng.permutation(g.num_nodes)  # new id of each old vertex
    inv = np.argsort(perm)
    deg = np.diff(g.index)[inv]
    idx = np.zeros(g.num_nodes + 1, np.int64)
    np.cumsum(deg, out=idx[1:])
    nb = np.concatenate([np.sort(perm[g.neighbors[g.index[o]:g.index[o + 1]]])
                         for o in inv]).astype(np.int32)
    h = hc.HostGraph.from_csr(idx, nb)
    sources = h.pick_sources(64, seed=4)
    monkeypatch.setenv("HC_FORCE_OFF64", "1")
    with hc.DeviceGraph(h) as dg:
        info = dg.info()
        assert info["offset_bytes"] == 8 and not info["degree_sorted"]
        new, old = _bfs_multi_both(hc, dg, sources)
    assert np.array_equal(new.values, old.values)
    for q in (0, 31, 63):
        assert np.array_equal(new.values[q], oracle.bfs_depth(idx, nb, int(sources[q]))), q


@pytest.mark.parametrize("shift", [9, 10, 12])
@pytest.mark.parametrize("kind,scale", [("kron", 14), ("uniform", 13)])
def test_bfs_multi_bitwise_column_blocks(hc, gpu, oracle, reset_knobs, shift, kind, scale):
    """Pull levels over column blocks (forced here: automatic only on flat
    graphs above 2^19 vertices), every level pulled, against the batched
    sweep and the oracle."""
    g = hc.HostGraph.generate(kind, scale, 16, weighted=False, seed=shift)
    sources = g.pick_sources(64, seed=shift)
    hc.set_tuning("msbfs.alpha", 1 << 20)
    hc.set_tuning("msbfs.col_shift", shift)
    with hc.DeviceGraph(g) as dg:
        new, old = _bfs_multi_both(hc, dg, sources)
        hc.set_tuning("msbfs.col_shift", 0)
        plain = hc.bfs_multi(dg, sources)
    assert new.stats["push_epochs"] == 0
    assert np.array_equal(new.values, old.values)
    assert np.array_equal(new.values, plain.values)
    for q in (0, 32, 63):
        assert np.array_equal(new.values[q], oracle.bfs_depth(g.index, g.neighbors,
                                                               int(sources[q]))), q


def test_checksum_kernels_match_restatement(hc, gpu):
    """hc_checksum_i32 (stride 1 and a column of a 64-wide matrix) and
    hc_checksum_multi against the restated sum, HC_INF_DIST read as -1 where
    asked; row counts not a multiple of the kernel's row groups."""
    from conftest import result_checksum
    rng = np.random.RandomState(4)
    for n in (1, 5, 1023, 70001):
        a = rng.randint(-3, 1 << 20, size=n).astype(np.int32)
        a[rng.rand(n) < 0.2] = hc.INF_DIST
        d = hc.DeviceArray.from_host(a)
        assert hc.checksum(d, n) == result_checksum(a), n
        m = rng.randint(0, 1 << 12, size=(n, hc.MULTI_WIDTH)).astype(np.int32)
        m[rng.rand(n, hc.MULTI_WIDTH) < 0.3] = hc.INF_DIST
        dm = hc.DeviceArray.from_host(m.ravel())
        cols = 64 if n % 2 else 37
        got = hc.checksum_multi(dm, n, cols, unreached_minus1=True)
        ref = [result_checksum(np.where(m[:, c] == hc.INF_DIST, -1, m[:, c])) for c in range(cols)]
        assert [int(x) for x in got] == ref, n
        raw = hc.checksum_multi(dm, n, 3)
        assert [int(x) for x in raw] == [result_checksum(m[:, c]) for c in range(3)]
        for c in (0, cols - 1):
            assert hc.checksum(dm, n, stride=hc.MULTI_WIDTH, column=c,
                               unreached_minus1=True) == ref[c], (n, c)
    empty = hc.DeviceArray(1, np.int32)
    assert hc.checksum(empty, 0) == 0
    with pytest.raises(hc.HcError):
        hc.checksum(empty, 1, stride=2)
