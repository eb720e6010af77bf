"""Pins of the oracle's shortest-path certificate (or_certify_paths) -- CPU.

The certificate is what the full-size GPU tests use to prove that every
distance and depth vector is exact, so it is pinned here against answers it
does not compute: it must accept the known-answer fixtures (scipy Dijkstra
and unweighted shortest paths), accept exactly one vector among every
candidate on tiny graphs (exhaustive enumeration: uniqueness), and reject
planted mistakes of each kind a kernel could make."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import golden_files, load_golden
from test_oracle import csr_from_edges, floyd

INF = 0x3FFFFFFF


@pytest.mark.parametrize("name", golden_files())
def test_certificate_accepts_fixtures(oracle, name):
    fx = load_golden(name)
    idx, nb, w = fx["index"], fx["neighbors"], fx["weights"]
    src = fx["sources"].astype(np.int32)
    depth = fx["depth"]
    D = np.ascontiguousarray(depth.T)
    assert list(oracle.certify_paths(idx, nb, None, src, D, -1)) == [0] * src.size
    if w.size and w.min() < 1:
        with pytest.raises(ValueError):
            oracle.certify_paths(idx, nb, w, src, np.ascontiguousarray(fx["sssp"].T))
        return
    D = np.ascontiguousarray(fx["sssp"].T)
    assert list(oracle.certify_paths(idx, nb, w, src, D)) == [0] * src.size


def test_certificate_unique_on_tiny_graphs(oracle):
    """Enumerate every vector over {0..7, unreached} on small graphs: the
    certificate accepts exactly the Floyd-Warshall distances."""
    cases = [
        (4, [(0, 1, 2), (1, 2, 1), (0, 2, 3), (2, 3, 2)]),
        (4, [(0, 1, 1), (2, 3, 1)]),                    # two components
        (5, [(0, 1, 1), (1, 2, 1), (2, 3, 1), (3, 4, 1), (4, 0, 1)]),
    ]
    vals = list(range(8)) + [INF]
    for n, edges in cases:
        idx, nb, w = csr_from_edges(n, edges)
        fw = floyd(idx, nb, w)
        cand = np.array(list(itertools.product(vals, repeat=n)), np.int32)  # [C, n]
        D = np.ascontiguousarray(cand.T)
        for s in range(n):
            bad = oracle.certify_paths(idx, nb, w, np.full(cand.shape[0], s, np.int32), D)
            ok = np.flatnonzero(bad == 0)
            truth = np.where(fw[s] >= INF, INF, fw[s])
            if truth[truth < INF].max() > 7:
                assert ok.size == 0
                continue
            assert ok.size == 1 and np.array_equal(cand[ok[0]], truth), (n, edges, s)


def test_certificate_unique_bfs_tiny(oracle):
    """Unit weights, -1 unreached: exactly the BFS depths are accepted."""
    n = 5
    idx, nb, _ = csr_from_edges(n, [(0, 1, 1), (1, 2, 1), (1, 3, 1)])
    vals = [-1, 0, 1, 2, 3, 4]
    cand = np.array(list(itertools.product(vals, repeat=n)), np.int32)
    D = np.ascontiguousarray(cand.T)
    for s in range(n):
        bad = oracle.certify_paths(idx, nb, None, np.full(cand.shape[0], s, np.int32), D, -1)
        ok = np.flatnonzero(bad == 0)
        assert ok.size == 1
        assert np.array_equal(cand[ok[0]], oracle.bfs_depth(idx, nb, s))


def test_certificate_rejects_planted_mistakes(oracle, hc):
    g = hc.HostGraph.generate("kron", 12, 16, weighted=True)
    idx, nb, w = g.index, g.neighbors, g.weights
    n = g.num_nodes
    src = int(g.pick_sources(1)[0])
    d = oracle.sssp_dijkstra(idx, nb, w, src)
    assert oracle.certify_sssp(idx, nb, w, src, d) == 0
    reached = np.flatnonzero((d < INF) & (np.arange(n) != src))
    unreached = np.flatnonzero(d == INF)
    rng = np.random.default_rng(7)
    u = int(rng.choice(reached))
    mistakes = []
    for delta in (+1, -1, +300):           # one distance off either way
        x = d.copy()
        x[u] += delta
        mistakes.append(x)
    x = d.copy()
    x[u] = INF                              # a reached vertex left at the sentinel
    mistakes.append(x)
    x = d.copy()
    x[src] = 1                              # source not at zero
    mistakes.append(x)
    x = d.copy()
    x[reached] += 1                         # uniform shift of everything but the source
    mistakes.append(x)
    x = oracle.sssp_dijkstra(idx, nb, w + 1, src)   # a relax with an extra unit
    mistakes.append(x)
    x = d.copy()
    x[u] = -5                               # negative garbage
    mistakes.append(x)
    if unreached.size:
        x = d.copy()
        x[unreached[0]] = 17                # unreachable vertex given a distance
        mistakes.append(x)
    # one Jacobi epoch short of the fixpoint: stale but consistent-looking
    y = np.full(n, INF, np.int32)
    y[src] = 0
    for _ in range(3):
        y, _ = oracle.sssp_jacobi_epoch(idx, nb, w, 0, n, y)
    if not np.array_equal(y, d):
        mistakes.append(y)
    D = np.ascontiguousarray(np.stack(mistakes, axis=1))
    bad = oracle.certify_paths(idx, nb, w, np.full(len(mistakes), src, np.int32), D)
    assert np.all(bad > 0), bad

    depth = oracle.bfs_depth(idx, nb, src)
    assert oracle.certify_bfs(idx, nb, src, depth) == 0
    wrong = []
    for delta in (+1, -1):
        x = depth.copy()
        x[u] += delta
        wrong.append(x)
    x = depth.copy()
    x[u] = -1
    wrong.append(x)
    x = depth.copy()
    x[x > 0] += 1
    wrong.append(x)
    D = np.ascontiguousarray(np.stack(wrong, axis=1))
    bad = oracle.certify_paths(idx, nb, None, np.full(len(wrong), src, np.int32), D, -1)
    assert np.all(bad > 0), bad


def test_certificate_random_perturbations(oracle):
    """Random graphs, random changes of 1-3 entries: every changed vector
    fails, every true vector passes (columns checked together)."""
    rng = np.random.default_rng(11)
    for trial in range(20):
        n = int(rng.integers(3, 40))
        m = int(rng.integers(n, 4 * n))
        edges = [(int(rng.integers(n)), int(rng.integers(n)), int(rng.integers(1, 9)))
                 for _ in range(m)]
        idx, nb, w = csr_from_edges(n, edges)
        fw = floyd(idx, nb, w)
        cols, srcs, want_ok = [], [], []
        for s in range(n):
            t = fw[s].astype(np.int32)
            cols.append(t)
            srcs.append(s)
            want_ok.append(True)
            x = t.copy()
            k = int(rng.integers(1, 4))
            for v in rng.choice(n, size=k, replace=False):
                choices = [c for c in (int(x[v]) + 1, int(x[v]) - 1, INF, 0, int(x[v]) + 9)
                           if c != x[v] and (c == INF or 0 <= c < INF)]
                x[v] = choices[int(rng.integers(len(choices)))]
            if not np.array_equal(x, t):
                cols.append(x)
                srcs.append(s)
                want_ok.append(False)
        D = np.ascontiguousarray(np.stack(cols, axis=1))
        bad = oracle.certify_paths(idx, nb, w, np.array(srcs, np.int32), D)
        assert list(bad == 0) == want_ok, trial


def test_certificate_layout_and_arguments(oracle, hc):
    """Vertex-major matrices with a leading dimension above k (the batched
    [N][64] layout) are read column by column; bad sources are refused."""
    g = hc.HostGraph.generate("uniform", 10, 8, weighted=True)
    idx, nb, w = g.index, g.neighbors, g.weights
    srcs = g.pick_sources(5)
    M = np.full((g.num_nodes, 64), INF, np.int32)
    for q, s in enumerate(srcs):
        M[:, q] = oracle.sssp_dijkstra(idx, nb, w, int(s))
    assert list(oracle.certify_paths(idx, nb, w, srcs, M)) == [0] * 5
    M[3, 2] += 1
    bad = oracle.certify_paths(idx, nb, w, srcs, M)
    assert bad[2] > 0 and bad[[0, 1, 3, 4]].sum() == 0
    with pytest.raises(ValueError):
        oracle.certify_paths(idx, nb, w, [g.num_nodes], M)
