"""ctypes binding of liboracle.so -- the CPU restatement of the reference's
epoch algorithms.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py.  The library under test never loads it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
INF = 0x3FFFFFFF

_P = C.c_void_p
_SIGS = {
    "or_version": ([], C.c_int32),
    "or_sssp_pull_serial": ([C.c_int32, _P, _P, _P, C.c_int32, _P], None),
    "or_sssp_dijkstra": ([C.c_int32, _P, _P, _P, C.c_int32, _P], None),
    "or_sssp_jacobi_epoch": ([C.c_int32, _P, _P, _P, C.c_int32, C.c_int32, _P, _P], C.c_int32),
    "or_sssp_pull_omp": ([C.c_int32, _P, _P, _P, _P, _P, _P, C.c_int32], C.c_double),
    "or_bfs_depth": ([C.c_int32, _P, _P, C.c_int32, _P], None),
    "or_verify_parents": ([C.c_int32, _P, _P, C.c_int32], C.c_int64),
    "or_bfs_pull_epoch": ([C.c_int32, _P, _P, _P, C.c_int32, C.c_int32, _P, _P], C.c_int32),
    "or_bfs_do": ([C.c_int32, _P, _P, C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_int32,
                   _P, C.c_int32], C.c_double),
    "or_pr_jacobi_f32": ([C.c_int32, _P, _P, C.c_float, C.c_int32, _P], None),
    "or_pr_jacobi_f64": ([C.c_int32, _P, _P, C.c_double, C.c_int32, _P], None),
    "or_pr_gauss_seidel_f32": ([C.c_int32, _P, _P, C.c_int32, _P], C.c_int32),
    "or_pr_jacobi_rule_f32": ([C.c_int32, _P, _P, C.c_float, C.c_int32, _P], C.c_int32),
    "or_pr_jacobi_rule_f32_order": ([C.c_int32, _P, _P, C.c_float, C.c_int32, C.c_int32, _P],
                                    C.c_int32),
    "or_equal_edge_ranges": ([C.c_int32, _P, C.c_int32, C.c_int32, _P], None),
    "or_certify_paths": ([C.c_int32, _P, _P, _P, C.c_int32, _P, _P, C.c_int64, C.c_int32,
                          _P], C.c_int64),
}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _csr(index, nbr, wt=None):
    index = np.ascontiguousarray(index, np.int64)
    nbr = np.ascontiguousarray(nbr, np.int32)
    wt = None if wt is None else np.ascontiguousarray(wt, np.int32)
    return index, nbr, wt


def sssp_pull_serial(index, nbr, wt, src):
    index, nbr, wt = _csr(index, nbr, wt)
    n = index.size - 1
    d = np.empty(n, np.int32)
    lib().or_sssp_pull_serial(n, _p(index), _p(nbr), _p(wt), src, _p(d))
    return d


def sssp_dijkstra(index, nbr, wt, src):
    index, nbr, wt = _csr(index, nbr, wt)
    n = index.size - 1
    d = np.empty(n, np.int32)
    lib().or_sssp_dijkstra(n, _p(index), _p(nbr), _p(wt), src, _p(d))
    return d


def sssp_jacobi_epoch(index, nbr, wt, start, end, din):
    index, nbr, wt = _csr(index, nbr, wt)
    din = np.ascontiguousarray(din, np.int32)
    out = np.empty_like(din)
    c = lib().or_sssp_jacobi_epoch(index.size - 1, _p(index), _p(nbr), _p(wt), start, end,
                                   _p(din), _p(out))
    return out, c


def sssp_pull_omp(index, nbr, wt, init, nthreads=0):
    index, nbr, wt = _csr(index, nbr, wt)
    init = np.ascontiguousarray(init, np.int32)
    d = np.empty_like(init)
    ep = np.zeros(1, np.int32)
    ms = lib().or_sssp_pull_omp(index.size - 1, _p(index), _p(nbr), _p(wt), _p(init),
                                _p(d), _p(ep), nthreads)
    return d, ms, int(ep[0])


def bfs_depth(index, nbr, src):
    index, nbr, _ = _csr(index, nbr)
    n = index.size - 1
    d = np.empty(n, np.int32)
    lib().or_bfs_depth(n, _p(index), _p(nbr), src, _p(d))
    return d


def verify_parents(depth, parents, src):
    depth = np.ascontiguousarray(depth, np.int32)
    parents = np.ascontiguousarray(parents, np.int32)
    return int(lib().or_verify_parents(depth.size, _p(depth), _p(parents), src))


def bfs_pull_epoch(index, nbr, parents, start, end, frontier):
    index, nbr, _ = _csr(index, nbr)
    parents = np.array(parents, np.int32, copy=True)
    frontier = np.ascontiguousarray(frontier, np.uint32)
    nxt = np.zeros_like(frontier)
    c = lib().or_bfs_pull_epoch(index.size - 1, _p(index), _p(nbr), _p(parents), start, end,
                                _p(frontier), _p(nxt))
    return parents, nxt, c


def bfs_do(index, nbr, src, alpha=15, beta=18, nthreads=0):
    index, nbr, _ = _csr(index, nbr)
    n = index.size - 1
    parents = np.empty(n, np.int32)
    dirs = np.zeros(max(n, 1) + 8, np.int8)
    nd = np.zeros(1, np.int32)
    ms = lib().or_bfs_do(n, _p(index), _p(nbr), src, alpha, beta, _p(parents), _p(dirs),
                         dirs.size, _p(nd), nthreads)
    return parents, ms, dirs[:int(nd[0])].copy()


def certify_paths(index, nbr, wt, sources, D, unreached=INF):
    """Certificate of distance (wt given) or depth (wt None) vectors:
    D is a C-contiguous [n, ld] int32 matrix whose column q claims the
    result from sources[q] (q < len(sources)).  Returns the per-column count
    of vertices that break the certificate (all zero iff every column is the
    exact answer; see or_certify_paths).  Raises ValueError for weights
    below 1 or sources out of range, where the certificate proves nothing."""
    index, nbr, wt = _csr(index, nbr, wt)
    n = index.size - 1
    D = np.ascontiguousarray(D, np.int32)
    if D.ndim == 1:
        D = D.reshape(n, 1)
    src = np.ascontiguousarray(np.atleast_1d(sources), np.int32)
    if D.shape[0] != n or D.shape[1] < src.size:
        raise ValueError(f"D has shape {D.shape}, need ({n}, >= {src.size})")
    bad = np.zeros(src.size, np.int64)
    r = lib().or_certify_paths(n, _p(index), _p(nbr), _p(wt), src.size, _p(src), _p(D),
                               D.shape[1], unreached, _p(bad))
    if r < 0:
        raise ValueError("certificate needs weights >= 1 and sources in range")
    return bad


def certify_sssp(index, nbr, wt, src, dist):
    """Violations of one SSSP distance vector (HC_INF_DIST = unreachable)."""
    return int(certify_paths(index, nbr, wt, [src], dist, INF)[0])


def certify_bfs(index, nbr, src, depth):
    """Violations of one BFS depth vector (-1 = unreachable)."""
    return int(certify_paths(index, nbr, None, [src], depth, -1)[0])


def pr_jacobi_f32(index, nbr, damping, iters):
    index, nbr, _ = _csr(index, nbr)
    n = index.size - 1
    s = np.empty(n, np.float32)
    lib().or_pr_jacobi_f32(n, _p(index), _p(nbr), damping, iters, _p(s))
    return s


def pr_jacobi_f64(index, nbr, damping, iters):
    index, nbr, _ = _csr(index, nbr)
    n = index.size - 1
    s = np.empty(n, np.float64)
    lib().or_pr_jacobi_f64(n, _p(index), _p(nbr), damping, iters, _p(s))
    return s


def pr_gauss_seidel_f32(index, nbr, max_iters):
    index, nbr, _ = _csr(index, nbr)
    n = index.size - 1
    s = np.empty(n, np.float32)
    it = lib().or_pr_gauss_seidel_f32(n, _p(index), _p(nbr), max_iters, _p(s))
    return s, int(it)


def pr_jacobi_rule_f32(index, nbr, damping, max_iters):
    """(scores, sweeps) of the Jacobi iteration under pr_pull_cpu's stopping
    rule (src/kernels/cpu/pr.cuh:165-181): a vertex takes its new score only
    when it moves by more than 5e-8; stop after a sweep that moved none."""
    index, nbr, _ = _csr(index, nbr)
    n = index.size - 1
    s = np.empty(n, np.float32)
    it = lib().or_pr_jacobi_rule_f32(n, _p(index), _p(nbr), damping, max_iters, _p(s))
    return s, int(it)


def pr_jacobi_rule_f32_reversed(index, nbr, damping, max_iters):
    """pr_jacobi_rule_f32 with each vertex's sum over its neighbours in reverse
    CSR order (another FP32 summation order of the same rule)."""
    index, nbr, _ = _csr(index, nbr)
    n = index.size - 1
    s = np.empty(n, np.float32)
    it = lib().or_pr_jacobi_rule_f32_order(n, _p(index), _p(nbr), damping, max_iters, 1,
                                           _p(s))
    return s, int(it)


def equal_edge_ranges(index, segments, chunk=1):
    index = np.ascontiguousarray(index, np.int64)
    out = np.zeros(segments + 1, np.int32)
    lib().or_equal_edge_ranges(index.size - 1, _p(index), segments, chunk, _p(out))
    return out
