"""Host side of the library (CPU only): the gapbs-style generator and
builder, the reference binary graph format, SourcePicker and the
equal-edge segmenter."""
from __future__ import annotations

import os

import numpy as np
import pytest


def check_csr_invariants(g, weighted=True):
    idx, nb = g.index, g.neighbors
    n = g.num_nodes
    assert idx[0] == 0 and idx[-1] == g.num_edges == nb.size
    deg = np.diff(idx)
    assert np.all(deg >= 0)
    rows = np.repeat(np.arange(n), deg)
    assert np.all((nb >= 0) & (nb < n))
    assert not np.any(nb == rows), "self-loop"
    # sorted rows without duplicates: strictly increasing within each row
    inner = np.diff(nb) > 0
    same_row = rows[1:] == rows[:-1]
    assert np.all(inner[same_row]), "row not strictly increasing"
    # symmetric, with equal weights in both directions
    key = rows.astype(np.int64) * n + nb
    rkey = nb.astype(np.int64) * n + rows
    o1, o2 = np.argsort(key), np.argsort(rkey)
    assert np.array_equal(key[o1], rkey[o2])
    if weighted:
        w = g.weights
        assert np.array_equal(w[o1], w[o2])
    return deg


@pytest.mark.parametrize("kind", ["kron", "uniform"])
def test_generator_invariants(hc, kind):
    g = hc.HostGraph.generate(kind, 12, 16, weighted=True)
    assert g.num_nodes == 4096
    deg = check_csr_invariants(g)
    assert np.all(deg[:-1] >= deg[1:]), "not relabelled by descending degree"
    w = g.weights
    assert w.min() >= 1 and w.max() <= 255
    # about 2*N*ef directed edges before duplicates/self-loops are dropped
    assert 0.6 * 2 * 4096 * 16 < g.num_edges <= 2 * 4096 * 16
    g2 = hc.HostGraph.generate(kind, 12, 16, weighted=True)
    assert np.array_equal(g.index, g2.index) and np.array_equal(g.neighbors, g2.neighbors)
    assert np.array_equal(g.weights, g2.weights)
    g3 = hc.HostGraph.generate(kind, 12, 16, weighted=True, seed=1)
    assert not np.array_equal(g.neighbors[:1000], g3.neighbors[:1000])


def test_kron_is_skewed_uniform_is_not(hc):
    k = hc.HostGraph.generate("kron", 14, 16, weighted=False)
    u = hc.HostGraph.generate("uniform", 14, 16, weighted=False)
    dk, du = k.degrees(), u.degrees()
    assert dk.max() > 20 * dk.mean() and (dk == 0).sum() > 0.1 * k.num_nodes
    assert du.max() < 4 * du.mean() and (du == 0).sum() < 0.01 * u.num_nodes
    assert not k.weighted


def test_builder_squish_and_relabel(hc):
    # duplicates keep the lightest weight, self-loops vanish, both directions
    src = [0, 1, 0, 2, 1, 2, 3, 0]
    dst = [1, 0, 1, 2, 2, 3, 3, 3]
    w = [5, 3, 9, 1, 4, 8, 2, 20]
    g = hc.HostGraph.from_edges(src, dst, 5, weights=w, symmetrize=True, relabel=False)
    assert list(g.index) == [0, 2, 4, 6, 8, 8]
    assert list(g.neighbors) == [1, 3, 0, 2, 1, 3, 0, 2]
    assert list(g.weights) == [3, 20, 3, 4, 4, 8, 20, 8]
    # relabel: degree descending, ties broken by larger old id first
    r = hc.HostGraph.from_edges(src, dst, 5, weights=w, symmetrize=True, relabel=True)
    check_csr_invariants(r)
    # old degrees 2,2,2,2,0 -> new order 3,2,1,0,4
    assert list(np.diff(r.index)) == [2, 2, 2, 2, 0]
    old_to_new = {3: 0, 2: 1, 1: 2, 0: 3, 4: 4}
    new_edges = {(old_to_new[a], old_to_new[b]) for a, b in [(0, 1), (0, 3), (1, 2), (2, 3)]}
    rows = np.repeat(np.arange(5), np.diff(r.index))
    got = {(int(a), int(b)) for a, b in zip(rows, r.neighbors) if a < b}
    assert got == {tuple(sorted(e)) for e in new_edges}


def test_builder_rejects_bad_input(hc):
    with pytest.raises(hc.HcError, match="outside"):
        hc.HostGraph.from_edges([0, 5], [1, 1], 3)
    with pytest.raises(hc.HcError, match="weight"):
        hc.HostGraph.from_edges([0], [1], 3, weights=[-1])
    with pytest.raises(hc.HcError):
        hc.HostGraph.generate("kron", 0, 16)
    g = hc.HostGraph.from_edges([], [], 4)
    assert g.num_edges == 0 and list(g.index) == [0] * 5


def test_file_roundtrip_and_reference_layout(hc, tmp_path):
    g = hc.HostGraph.generate("kron", 10, 8, weighted=True)
    p = str(tmp_path / "g.wsg")
    g.write(p, weighted=True)
    raw = open(p, "rb").read()
    n = np.frombuffer(raw[:4], np.int32)[0]
    m = np.frombuffer(raw[4:12], np.int64)[0]
    assert n == g.num_nodes and m == g.num_edges
    assert len(raw) == 12 + 8 * (n + 1) + 8 * m
    rec = np.frombuffer(raw[12 + 8 * (n + 1):], dtype=[("v", "<i4"), ("w", "<f4")])
    assert np.array_equal(rec["v"], g.neighbors) and np.array_equal(rec["w"], g.weights)
    h = hc.HostGraph.read(p, weighted=True)
    assert np.array_equal(h.index, g.index) and np.array_equal(h.weights, g.weights)
    # unweighted file as the reference writes it: 8 bytes per edge from a
    # 4-byte array (trailing bytes after the ids) -- the reader ignores them
    q = str(tmp_path / "g.sg")
    with open(q, "wb") as f:
        f.write(raw[:12 + 8 * (n + 1)])
        f.write(g.neighbors.tobytes())
        f.write(b"\x00" * (4 * m))
    u = hc.HostGraph.read(q, weighted=False)
    assert np.array_equal(u.neighbors, g.neighbors) and not u.weighted
    u.write(str(tmp_path / "h.sg"), weighted=False)
    assert os.path.getsize(tmp_path / "h.sg") == 12 + 8 * (n + 1) + 4 * m


def test_file_validation(hc, tmp_path):
    def write(name, n, m, index, payload):
        p = str(tmp_path / name)
        with open(p, "wb") as f:
            f.write(np.int32(n).tobytes() + np.int64(m).tobytes())
            f.write(np.asarray(index, np.int64).tobytes())
            f.write(payload)
        return p

    with pytest.raises(hc.HcError, match="cannot open"):
        hc.HostGraph.read(str(tmp_path / "missing.wsg"), True)
    with pytest.raises(hc.HcError, match="bad header"):
        hc.HostGraph.read(write("a.sg", -1, 0, [], b""), False)
    with pytest.raises(hc.HcError, match="bytes"):
        hc.HostGraph.read(write("b.sg", 2, 4, [0, 2, 4], b"\0" * 8), False)
    with pytest.raises(hc.HcError, match="monotone|index"):
        hc.HostGraph.read(write("c.sg", 2, 2, [0, 3, 2], np.zeros(2, np.int32).tobytes()), False)
    with pytest.raises(hc.HcError, match="out of range"):
        hc.HostGraph.read(write("d.sg", 2, 2, [0, 1, 2], np.array([1, 7], np.int32).tobytes()), False)
    rec = np.array([(1, 0.5), (0, 0.5)], dtype=[("v", "<i4"), ("w", "<f4")])
    with pytest.raises(hc.HcError, match="integer"):
        hc.HostGraph.read(write("e.wsg", 2, 2, [0, 1, 2], rec.tobytes()), True)


def _lemire(raw, rng_range):
    """libstdc++ (GCC >= 11) uniform_int_distribution for a 32-bit engine:
    Lemire's nearly-divisionless method."""
    i = 0
    while True:
        x = int(raw[i]); i += 1
        prod = x * rng_range
        low = prod & 0xFFFFFFFF
        if low < rng_range:
            thresh = (-rng_range % (1 << 32)) % rng_range
            while low < thresh:
                x = int(raw[i]); i += 1
                prod = x * rng_range
                low = prod & 0xFFFFFFFF
        return prod >> 32, i


def test_source_picker_matches_mt19937(hc):
    """SourcePicker (src/graph.cuh:106-141): mt19937 seeded 37525 (numpy's
    legacy MT19937 seeding is the same init_genrand), uniform ids by
    libstdc++'s mapping, degree-0 vertices rejected."""
    g = hc.HostGraph.generate("kron", 12, 4, weighted=False)
    deg = g.degrees()
    assert (deg == 0).sum() > 0
    got = g.pick_sources(40)
    rs = np.random.RandomState(37525)
    raw = [rs.randint(0, 2**32, dtype=np.uint64) for _ in range(4000)]
    pos, exp = 0, []
    while len(exp) < 40:
        v, used = _lemire(raw[pos:], g.num_nodes)
        pos += used
        if deg[v] > 0:
            exp.append(v)
    assert list(got) == exp
    assert np.array_equal(g.pick_sources(40), got)
    with pytest.raises(hc.HcError):
        hc.HostGraph.from_edges([], [], 3).pick_sources(1)


def test_equal_edge_ranges(hc, oracle):
    g = hc.HostGraph.generate("kron", 13, 16, weighted=False)
    idx = g.index
    maxdeg = int(np.diff(idx).max())
    for S, chunk in [(24, 1), (24, 32), (7, 1), (1, 1), (100, 64)]:
        r = g.equal_edge_ranges(S, chunk)
        assert r[0] == 0 and r[-1] == g.num_nodes and np.all(np.diff(r) >= 0)
        edges = np.diff(idx[r])
        assert edges.sum() == g.num_edges
        if chunk == 1:
            if S > 1:
                assert edges[:-1].max() <= g.num_edges // S + maxdeg
            ref = oracle.equal_edge_ranges(idx, S, chunk)
            # identical wherever the reference sets a boundary
            set_ = ref >= 0
            assert np.array_equal(r[set_][:-1], ref[set_][:-1])
        else:
            assert all(x % chunk == 0 or x == g.num_nodes for x in r)
