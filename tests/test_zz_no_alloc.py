"""GPU test of hc_bfs_many_noalloc, the many-graph BFS that allocates no
device memory: through the C ABI it must give the same depths as
hc_bfs_many (and valid parents), on two calls in a row, the second with a
smaller batch, while the GPU's free memory does not move."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


def _batch_graphs(hc, seed, count, big_scale):
    """`count` Kronecker graphs of scale 8..13 plus one of `big_scale`, whose
    bitmaps (above 2^17 vertices) live in the scratch slots."""
    rng = np.random.RandomState(seed)
    gs = [hc.HostGraph.generate("kron", int(rng.randint(8, 14)), int(rng.randint(8, 25)),
                                weighted=False, seed=seed + i) for i in range(count)]
    gs.insert(count // 2, hc.HostGraph.generate("kron", big_scale, 8, weighted=False,
                                                 seed=seed + 999))
    return gs


def _compare(hc, oracle, graphs):
    srcs = [int(g.pick_sources(1)[0]) for g in graphs]
    with hc.GraphBatch(graphs) as b:
        d0, p0, st0 = b.bfs(srcs, noalloc=True)
        # The first call in a process loads the library's code object for
        # this kernel, which holds the fixed scratch; a repeat call must not
        # change the free device memory (one bitmap slot of the large graph
        # alone would be >= 96 KiB per workgroup).
        hc.synchronize()
        free0, _ = hc.mem_get_info()
        d2, _, _ = b.bfs(srcs, noalloc=True)
        hc.synchronize()
        free1, _ = hc.mem_get_info()
        d1, p1, st1 = b.bfs(srcs)
    assert free0 - free1 < (1 << 20), (free0, free1)
    assert all(np.array_equal(x, y) for x, y in zip(d0, d2))
    assert (st0["push_epochs"], st0["pull_epochs"]) == (st1["push_epochs"],
                                                        st1["pull_epochs"])
    for i, g in enumerate(graphs):
        assert np.array_equal(d0[i], d1[i]), i
        want = oracle.bfs_depth(g.index, g.neighbors, srcs[i])
        assert np.array_equal(d0[i], want), i
        assert oracle.verify_parents(want, p0[i], srcs[i]) == 0, i


def test_noalloc_matches_allocating_entry(hc, gpu, oracle):
    # First call: 50 graphs with a scale-21 one (2^21 vertices: 768 KiB of
    # bitmaps per workgroup, so only 42 workgroups fit the 32 MiB scratch and
    # the grid shrinks below the graph count).  Second call: a smaller batch
    # (12 graphs, largest scale 18) over the scratch the first call left.
    print("noalloc: call 1 (50 + kron-21)", flush=True)
    _compare(hc, oracle, _batch_graphs(hc, 500, 49, 21))
    print("noalloc: call 2 (12 + kron-18)", flush=True)
    _compare(hc, oracle, _batch_graphs(hc, 900, 11, 18))


def test_noalloc_rejects_batch_above_scratch(hc, gpu):
    g = hc.HostGraph.generate("kron", 4, 4, weighted=False, seed=3)
    count = (1 << 16) + 1
    arr = (hc.HostGraphStruct * count)(*([g._s] * count))
    h = C.c_void_p()
    hc.check(hc.lib().hc_graph_batch_upload(arr, count, C.byref(h)))
    try:
        src = np.full(count, int(g.pick_sources(1)[0]), np.int32)
        d = hc.DeviceArray(count * g.num_nodes, np.int32)
        st = hc.RunStats()
        rc = hc.lib().hc_bfs_many_noalloc(h, src.ctypes.data, d.ptr, None, 15, 18,
                                          C.byref(st), None)
        assert rc == -1, rc  # HC_ERR_INVALID
        assert "65536" in hc.lib().hc_last_error().decode()
        d.free()
    finally:
        hc.lib().hc_graph_batch_free(h)
