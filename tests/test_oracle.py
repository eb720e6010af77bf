"""Pins of the CPU oracle (oracle/oracle.c) -- CPU only.

Each oracle function is checked against answers it does not compute itself:
the known-answer fixtures (scipy Dijkstra / unweighted shortest paths and a
numpy PageRank iteration plus dense solve, tests/golden/make_golden.py),
brute force on tiny graphs (Floyd-Warshall, exhaustive first-neighbour
scans), closed forms (paths, stars, cycles, complete and regular graphs) and
invariants.  A planted mistake (an extra weight unit, a dropped damping
term, a wrong parent) must fail the same pins."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_files, load_golden

INF = 0x3FFFFFFF


def csr_from_edges(n, edges, weighted=True):
    """Symmetric CSR from (u, v, w) triples (rows sorted, lightest duplicate
    kept, no self-loops) built with plain Python, independent of the
    library's builder."""
    rows = [dict() for _ in range(n)]
    for u, v, w in edges:
        if u == v:
            continue
        for a, b in ((u, v), (v, u)):
            rows[a][b] = min(w, rows[a].get(b, w))
    index = np.zeros(n + 1, np.int64)
    nbr, wt = [], []
    for u in range(n):
        for v in sorted(rows[u]):
            nbr.append(v)
            wt.append(rows[u][v])
        index[u + 1] = len(nbr)
    return index, np.array(nbr, np.int32), np.array(wt, np.int32)


def floyd(index, nbr, wt):
    n = index.size - 1
    d = np.full((n, n), np.inf)
    np.fill_diagonal(d, 0)
    for u in range(n):
        for i in range(index[u], index[u + 1]):
            v = nbr[i]
            d[v, u] = min(d[v, u], wt[i])  # edge v -> u (pull row u lists v)
    for k in range(n):
        d = np.minimum(d, d[:, [k]] + d[[k], :])
    return np.where(np.isinf(d), INF, d).astype(np.int64)


# ---------------------------------------------------------------------------
# Fixtures
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", golden_files())
def test_sssp_oracles_match_fixtures(oracle, name):
    fx = load_golden(name)
    idx, nb, w = fx["index"], fx["neighbors"], fx["weights"]
    for i, s in enumerate(fx["sources"]):
        s = int(s)
        want = fx["sssp"][i]
        assert np.array_equal(oracle.sssp_dijkstra(idx, nb, w, s), want)
        assert np.array_equal(oracle.sssp_pull_serial(idx, nb, w, s), want)
        init = np.full(idx.size - 1, INF, np.int32)
        init[s] = 0
        d, _, epochs = oracle.sssp_pull_omp(idx, nb, w, init, 4)
        assert np.array_equal(d, want) and epochs >= 1


@pytest.mark.parametrize("name", golden_files())
def test_bfs_oracles_match_fixtures(oracle, name):
    fx = load_golden(name)
    idx, nb = fx["index"], fx["neighbors"]
    for i, s in enumerate(fx["sources"]):
        s = int(s)
        depth = oracle.bfs_depth(idx, nb, s)
        assert np.array_equal(depth, fx["depth"][i])
        par, _, dirs = oracle.bfs_do(idx, nb, s, 15, 18, 4)
        assert oracle.verify_parents(depth, par, s) == 0
        # every reached level is one epoch: epochs = eccentricity + 1
        assert len(dirs) == depth.max() + 1


@pytest.mark.parametrize("name", golden_files())
def test_pagerank_oracles_match_fixtures(oracle, name):
    fx = load_golden(name)
    idx, nb = fx["index"], fx["neighbors"]
    p64 = oracle.pr_jacobi_f64(idx, nb, 0.85, 20)
    assert np.allclose(p64, fx["pr20"], rtol=0, atol=1e-15)
    p32 = oracle.pr_jacobi_f32(idx, nb, 0.85, 20)
    assert np.abs(p32 - fx["pr20"]).sum() < 1e-6
    if fx["prfix"].size:
        pinf = oracle.pr_jacobi_f64(idx, nb, 0.85, 400)
        assert np.allclose(pinf, fx["prfix"], rtol=0, atol=1e-12)


# ---------------------------------------------------------------------------
# Brute force and closed forms
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_sssp_vs_floyd_warshall(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    m = int(rng.integers(0, 3 * n))
    edges = [(int(rng.integers(n)), int(rng.integers(n)), int(rng.integers(0, 256)))
             for _ in range(m)]
    idx, nb, w = csr_from_edges(n, edges)
    fw = floyd(idx, nb, w)
    for s in range(0, n, max(1, n // 5)):
        assert np.array_equal(oracle.sssp_dijkstra(idx, nb, w, s), fw[s])
        assert np.array_equal(oracle.sssp_pull_serial(idx, nb, w, s), fw[s])


def test_sssp_closed_forms(oracle):
    n = 50
    # path with weight i+1 on edge (i, i+1): prefix sums
    idx, nb, w = csr_from_edges(n, [(i, i + 1, i + 1) for i in range(n - 1)])
    d = oracle.sssp_dijkstra(idx, nb, w, 0)
    assert np.array_equal(d, np.cumsum(np.r_[0, np.arange(1, n)]))
    # star: center 0, leaf i weight i; from leaf 3 the distance to leaf j is 3+j
    idx, nb, w = csr_from_edges(10, [(0, i, i) for i in range(1, 10)])
    d = oracle.sssp_pull_serial(idx, nb, w, 3)
    assert d[0] == 3 and all(d[j] == 3 + j for j in range(1, 10) if j != 3) and d[3] == 0
    # disconnected vertex stays at the sentinel; saturation: INF + w == INF
    idx, nb, w = csr_from_edges(4, [(0, 1, 255), (1, 2, 255)])
    d = oracle.sssp_dijkstra(idx, nb, w, 0)
    assert list(d) == [0, 255, 510, INF]


def test_bfs_closed_forms(oracle):
    n = 13
    idx, nb, _ = csr_from_edges(n, [(i, (i + 1) % n, 1) for i in range(n)])
    d = oracle.bfs_depth(idx, nb, 0)
    assert list(d) == [min(k, n - k) for k in range(n)]
    idx, nb, _ = csr_from_edges(7, [(i, j, 1) for i in range(7) for j in range(i + 1, 7)])
    assert list(oracle.bfs_depth(idx, nb, 2)) == [1, 1, 0, 1, 1, 1, 1]
    idx, nb, _ = csr_from_edges(6, [(0, 1, 1), (2, 3, 1)])
    assert list(oracle.bfs_depth(idx, nb, 0)) == [0, 1, -1, -1, -1, -1]


def test_pagerank_closed_forms(oracle):
    # regular graphs (cycle, complete): 1/N is the fixed point
    for edges, n in ([(i, (i + 1) % 17, 1) for i in range(17)], 17), \
                    ([(i, j, 1) for i in range(6) for j in range(i + 1, 6)], 6):
        idx, nb, _ = csr_from_edges(n, edges)
        p = oracle.pr_jacobi_f64(idx, nb, 0.85, 30)
        assert np.allclose(p, 1.0 / n, atol=1e-15)
    # star with k leaves: c = b + d*k*l, l = b + d*c/k  (dangling mass 0)
    k = 7
    n = k + 1
    idx, nb, _ = csr_from_edges(n, [(0, i, 1) for i in range(1, n)])
    p = oracle.pr_jacobi_f64(idx, nb, 0.85, 500)
    b, d = 0.15 / n, 0.85
    c = (b + d * k * b) / (1 - d * d)
    leaf = b + d * c / k
    assert abs(p[0] - c) < 1e-14 and np.allclose(p[1:], leaf, atol=1e-14)


def test_pagerank_gauss_seidel_reaches_jacobi_fixed_point(oracle, hc):
    g = hc.HostGraph.generate("kron", 9, 8, weighted=False)
    idx, nb = g.index, g.neighbors
    gs, sweeps = oracle.pr_gauss_seidel_f32(idx, nb, 1000)
    fix = oracle.pr_jacobi_f64(idx, nb, 0.85, 300)
    assert sweeps < 1000
    # stopping rule is a per-sweep change of 5e-8, so the error is larger
    assert np.abs(gs - fix).max() < 1e-4 * fix.max()


def test_pagerank_per_vertex_rule(oracle, hc):
    """pr_pull_cpu's stopping rule in Jacobi form: on a cycle every score is
    already the fixed point 1/N (one sweep, nothing moves); on a Kronecker
    graph the first sweep is the plain Jacobi sweep wherever a vertex moved
    by more than 5e-8 and the start value elsewhere, and at the stop no
    vertex is more than 5e-8 (plus FP32 rounding) from its FP64 update."""
    n = 1000
    idx = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    nb = np.array([[(u - 1) % n, (u + 1) % n] for u in range(n)], np.int32).ravel()
    s, sweeps = oracle.pr_jacobi_rule_f32(idx, nb, 0.85, 201)
    assert sweeps == 1 and np.all(s == np.float32(1.0 / n))

    g = hc.HostGraph.generate("kron", 14, 16)
    idx, nb = g.index, g.neighbors
    n = idx.size - 1
    init = np.float32(1.0) / np.float32(n)
    one, sweeps = oracle.pr_jacobi_rule_f32(idx, nb, 0.85, 1)
    j1 = oracle.pr_jacobi_f32(idx, nb, 0.85, 1)
    moved = np.abs(j1 - init) > np.float32(5e-8)
    assert sweeps == 1 and 0 < (~moved).sum() < n
    assert np.array_equal(one, np.where(moved, j1, init))

    s, sweeps = oracle.pr_jacobi_rule_f32(idx, nb, 0.85, 201)
    assert 1 < sweeps < 201
    deg = np.diff(idx).astype(np.float64)
    contrib = np.where(deg > 0, s.astype(np.float64) / np.maximum(deg, 1.0), 0.0)
    rows = np.repeat(np.arange(n), np.diff(idx))
    upd = 0.15 / n + 0.85 * np.bincount(rows, weights=contrib[nb], minlength=n)
    assert np.abs(upd - s).max() <= 5e-8 + 1e-10
    # The last sweep moved nothing; the one before it moved vertices.
    s1, _ = oracle.pr_jacobi_rule_f32(idx, nb, 0.85, sweeps - 1)
    s2, _ = oracle.pr_jacobi_rule_f32(idx, nb, 0.85, sweeps - 2)
    assert np.array_equal(s1, s) and not np.array_equal(s2, s)


def test_pagerank_per_vertex_rule_reversed_sums(oracle, hc):
    """The rule with each FP32 sum taken in reverse neighbour order: its
    first sweep equals a sequential FP32 loop over the reversed rows, which
    differs from the forward order at some vertices; at the stop the same
    5e-8 invariant holds."""
    g = hc.HostGraph.generate("kron", 9, 16)
    idx, nb = g.index, g.neighbors
    n = idx.size - 1
    deg = np.diff(idx)
    cur = np.full(n, np.float32(1.0) / np.float32(n), np.float32)
    want = np.empty(n, np.float32)
    for u in range(n):
        t = np.float32(0.0)
        for i in range(idx[u + 1] - 1, idx[u] - 1, -1):
            v = nb[i]
            t = np.float32(t + np.float32(cur[v] / np.float32(deg[v])))
        base = np.float32(np.float32(1.0) - np.float32(0.85)) / np.float32(n)
        ns = np.float32(base + np.float32(np.float32(0.85) * t))
        want[u] = ns if abs(np.float32(ns - cur[u])) > np.float32(5e-8) else cur[u]
    one, sweeps = oracle.pr_jacobi_rule_f32_reversed(idx, nb, 0.85, 1)
    assert sweeps == 1 and np.array_equal(one, want)
    fwd = oracle.pr_jacobi_f32(idx, nb, 0.85, 1)
    rev = oracle.pr_jacobi_rule_f32_reversed(idx, nb, 0.85, 1)[0]
    assert not np.array_equal(fwd, rev)  # the order matters in FP32

    s, sweeps = oracle.pr_jacobi_rule_f32_reversed(idx, nb, 0.85, 201)
    assert 1 < sweeps < 201
    contrib = np.where(deg > 0, s.astype(np.float64) / np.maximum(deg, 1), 0.0)
    rows = np.repeat(np.arange(n), deg)
    upd = 0.15 / n + 0.85 * np.bincount(rows, weights=contrib[nb], minlength=n)
    assert np.abs(upd - s).max() <= 5e-8 + 1e-10


def test_bfs_pull_epoch_first_neighbour(oracle, hc):
    """epoch_bfs_pull_cpu restatement vs an exhaustive scan."""
    g = hc.HostGraph.generate("kron", 10, 8, weighted=False)
    idx, nb = g.index, g.neighbors
    n = g.num_nodes
    rng = np.random.default_rng(3)
    parents = np.where(rng.random(n) < 0.3, rng.integers(0, n, n), -1).astype(np.int32)
    front_set = rng.random(n) < 0.1
    words = np.zeros((n + 31) // 32, np.uint32)
    for v in np.flatnonzero(front_set):
        words[v >> 5] |= np.uint32(1 << (v & 31))
    par, nxt, cnt = oracle.bfs_pull_epoch(idx, nb, parents, 100, 900, words)
    exp = parents.copy()
    exp_cnt = 0
    for u in range(100, 900):
        if parents[u] != -1:
            continue
        hits = [v for v in nb[idx[u]:idx[u + 1]] if front_set[v]]
        if hits:
            exp[u] = hits[0]
            exp_cnt += 1
            assert (nxt[u >> 5] >> (u & 31)) & 1
    assert cnt == exp_cnt and np.array_equal(par, exp)


def test_jacobi_epoch_brackets(oracle, hc):
    """One Jacobi epoch from d0 is >= the fixpoint and <= d0; iterating it
    reaches the Dijkstra fixpoint."""
    g = hc.HostGraph.generate("uniform", 9, 6, weighted=True)
    idx, nb, w = g.index, g.neighbors, g.weights
    d = np.full(g.num_nodes, INF, np.int32)
    d[5] = 0
    fix = oracle.sssp_dijkstra(idx, nb, w, 5)
    for _ in range(g.num_nodes):
        nd, c = oracle.sssp_jacobi_epoch(idx, nb, w, 0, g.num_nodes, d)
        assert np.all(nd <= d) and np.all(nd >= fix)
        d = nd
        if c == 0:
            break
    assert np.array_equal(d, fix)


# ---------------------------------------------------------------------------
# Planted mistakes fail the pins
# ---------------------------------------------------------------------------

def test_planted_mistakes_are_caught(oracle):
    fx = load_golden("kron10.npz")
    idx, nb, w = fx["index"], fx["neighbors"], fx["weights"]
    s = int(fx["sources"][0])
    # an extra unit on every weight (a mis-restated relax) changes distances
    assert not np.array_equal(oracle.sssp_dijkstra(idx, nb, w + 1, s), fx["sssp"][0])
    # a dropped damping factor changes PageRank
    assert np.abs(oracle.pr_jacobi_f64(idx, nb, 1.0, 20) - fx["pr20"]).sum() > 1e-3
    # a wrong parent / source / unreached flag is reported
    depth = fx["depth"][0]
    par, _, _ = oracle.bfs_do(idx, nb, s, 15, 18, 2)
    bad = par.copy()
    u = int(np.flatnonzero(depth == 2)[0])
    bad[u] = s  # depth 0 parent for a depth-2 vertex
    assert oracle.verify_parents(depth, bad, s) == 1
    bad = par.copy()
    bad[s] = int(np.flatnonzero(depth == 1)[0])
    assert oracle.verify_parents(depth, bad, s) == 1
    bad = par.copy()
    bad[int(np.flatnonzero(depth == 1)[0])] = -1
    assert oracle.verify_parents(depth, bad, s) == 1


def test_equal_edge_ranges_restatement(oracle, hc):
    """Oracle restatement of compute_equal_edge_ranges vs a direct Python
    transcription of the same loop (src/graph.cuh:156-201)."""
    g = hc.HostGraph.generate("kron", 12, 8, weighted=False)
    idx = g.index
    for S, chunk in [(24, 1), (8, 32), (5, 7), (1, 1)]:
        got = oracle.equal_edge_ranges(idx, S, chunk)
        n = idx.size - 1
        avg = int(idx[-1]) // S
        exp = [-1] * (S + 1)
        exp[0] = 0
        end = seg = sdeg = 0
        while end != n:
            sdeg += int(idx[end + 1] - idx[end])
            if sdeg >= avg:
                if seg + 1 <= S:
                    exp[seg + 1] = end
                sdeg = 0
                seg += 1
            end += 1
        if seg != S and seg + 1 <= S:
            exp[seg + 1] = end
        for i in range(1, S):
            if exp[i] >= 0 and exp[i] % chunk:
                exp[i] = (exp[i] // chunk + (exp[i] % chunk >= chunk // 2)) * chunk
        assert list(got) == exp


# ---------------------------------------------------------------------------
# bfs_do_cpu's push/pull switching rule, pinned by hand-derived sequences
# ---------------------------------------------------------------------------
# Each sequence below was worked out by hand from the rule of
# src/kernels/cpu/bfs.cuh:63-101 (integer division; pull while
# num_edges > edges_to_check / alpha; stay in pull while n > prev or
# n > N / beta; num_edges = 1 after a pull phase; a push subtracts the
# scout count before it runs), not by running a loop.  0 = push, 1 = pull.
#
# Star, centre + 8 leaves (N 9, m 16), from the centre: deg 8 > 16/15 = 1
#   -> pull finds 8 (8 > 1: stay) -> pull finds 0 (0 <= 8, 0 <= 9/18 = 0).
# Star, centre + 9 leaves (N 10, m 18), from a leaf: 1 > 18/15 = 1 is false
#   -> push (scout 9) -> 9 > 17/15 -> pull 8 -> pull 0.
# K_6 (m 30) from any vertex: 5 > 30/15 = 2 -> pull 5 -> pull 0.
# Cycle C_12 from 0: 2 > 24/15 = 1 -> pull; N/18 = 0 keeps every pull that
#   finds a vertex, so 7 pulls (levels 1..6, then an empty one).
# Path P_50 from vertex 0 (m 98): pushes while 2 > e/15 fails; after push j
#   e = 99 - 2j, first true at j = 35 (e = 29); the pull finds 1 node
#   (1 <= 1 and 1 <= 50/18 = 2: leave); then 1 > 29/15 = 1 fails -> push
#   (e 28), 2 > 1 -> pull, ... alternating for levels 37..48; push to 49
#   (scout 1: 1 > 22/15 = 1 fails), and a last empty push.
# Cycle C_40 from 0 (m 80, scout 4 per level after the first): pushes at
#   e = 80, 78, 74, 70, 66, 62 (4 > e/15 false down to e = 62); at e = 58
#   4 > 3 -> pull finds 2 (2 > 2 false, 2 > 40/18 = 2 false: leave); then
#   push / pull alternate (e 57..52), the push at level 20 finds one vertex
#   (scout 2 > 51/15 = 3 false) and the last push finds none.
BFS_DO_SEQUENCES = {
    "star8_centre": [1, 1],
    "star9_leaf": [0, 1, 1],
    "k6": [1, 1],
    "cycle12": [1] * 7,
    "path50": [0] * 35 + [1] + [0, 1] * 6 + [0, 0],
    "cycle40": [0] * 6 + [1] + [0, 1] * 6 + [0, 0],
}


def _do_case(name):
    if name == "star8_centre":
        return csr_from_edges(9, [(0, i, 1) for i in range(1, 9)]), 0
    if name == "star9_leaf":
        return csr_from_edges(10, [(0, i, 1) for i in range(1, 10)]), 4
    if name == "k6":
        return csr_from_edges(6, [(i, j, 1) for i in range(6) for j in range(i + 1, 6)]), 3
    if name == "cycle12":
        return csr_from_edges(12, [(i, (i + 1) % 12, 1) for i in range(12)]), 0
    if name == "path50":
        return csr_from_edges(50, [(i, i + 1, 1) for i in range(49)]), 0
    if name == "cycle40":
        return csr_from_edges(40, [(i, (i + 1) % 40, 1) for i in range(40)]), 0
    raise KeyError(name)


@pytest.mark.parametrize("name", sorted(BFS_DO_SEQUENCES))
def test_bfs_do_switching_sequence(oracle, name):
    (idx, nb, _), src = _do_case(name)
    for threads in (1, 4):
        par, _, dirs = oracle.bfs_do(idx, nb, src, 15, 18, threads)
        assert list(dirs) == BFS_DO_SEQUENCES[name], name
        assert oracle.verify_parents(oracle.bfs_depth(idx, nb, src), par, src) == 0


def test_bfs_do_switching_sequence_detects_slips(oracle):
    """The same hand-derived sequences reject plausible slips of the rule:
    alpha and beta swapped, and a push->pull test of >= instead of >
    (equivalently, a threshold one lower)."""
    (idx, nb, _), src = _do_case("path50")
    _, _, swapped = oracle.bfs_do(idx, nb, src, 18, 15, 1)
    assert list(swapped) != BFS_DO_SEQUENCES["path50"]
    (idx, nb, _), src = _do_case("cycle40")
    _, _, swapped = oracle.bfs_do(idx, nb, src, 18, 15, 1)
    assert list(swapped) != BFS_DO_SEQUENCES["cycle40"]
    # beta 17 keeps the first pull phase of C_40 going (2 > 40/17 = 2 is
    # still false) but beta 21 (40/21 = 1) does: the sequence changes
    _, _, b21 = oracle.bfs_do(idx, nb, src, 15, 21, 1)
    assert list(b21) != BFS_DO_SEQUENCES["cycle40"]
    # alpha 14 moves the first switch of P_50 (e/14 < 2 at e = 27, not 29)
    (idx, nb, _), src = _do_case("path50")
    _, _, a14 = oracle.bfs_do(idx, nb, src, 14, 18, 1)
    assert list(a14) != BFS_DO_SEQUENCES["path50"]
