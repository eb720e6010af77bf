"""Python binding of libhcamd.so, the gfx950 graph epoch library.

Thin ctypes layer over the C ABI declared in include/hc_capi.h: host CSR
graphs (generator, reference binary files, SourcePicker, equal-edge
segments), device-resident graphs in HBM, the epoch kernels and the full
BFS / SSSP / PageRank runs.  There is no fallback path: if the shared
library is missing or a call fails, an exception is raised.

Reference counterparts (chrisliu/hetero-compute): src/graph.cuh (CSR, loader,
SourcePicker, compute_equal_edge_ranges), src/kernels/gpu/{sssp,bfs,pr}.cuh
(drivers and epoch kernels), src/kernels/cpu/bfs.cuh:48-114 (bfs_do_cpu).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhcamd.so")

INF_DIST = 0x3FFFFFFF
INVALID_NODE = -1

MAP_THREAD, MAP_WAVE, MAP_BLOCK, MAP_BUCKETED = 0, 1, 2, 3

HC_OK = 0
ERR_NAMES = {
    -1: "HC_ERR_INVALID", -2: "HC_ERR_HIP", -3: "HC_ERR_OOM", -4: "HC_ERR_IO",
    -5: "HC_ERR_UNSUPPORTED", -6: "HC_ERR_BUSY", -7: "HC_ERR_NOT_FOUND",
    -8: "HC_ERR_NO_DEVICE",
}


class HcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class HostGraphStruct(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("num_edges", C.c_int64),
                ("index", C.POINTER(C.c_int64)),
                ("neighbors", C.POINTER(C.c_int32)),
                ("weights", C.POINTER(C.c_int32))]


class GraphInfo(C.Structure):
    _fields_ = [("num_nodes", C.c_int32), ("num_edges", C.c_int64),
                ("start_id", C.c_int32), ("end_id", C.c_int32),
                ("degree_sorted", C.c_int32), ("weighted", C.c_int32),
                ("offset_bytes", C.c_int32), ("num_nonzero", C.c_int32),
                ("max_degree", C.c_int64), ("device_bytes", C.c_size_t)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("arch", C.c_char * 64),
                ("compute_units", C.c_int32), ("wavefront_size", C.c_int32),
                ("hbm_bytes", C.c_size_t), ("lds_bytes_per_block", C.c_int32)]


class RunStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("epochs", C.c_int32),
                ("push_epochs", C.c_int32), ("pull_epochs", C.c_int32),
                ("launches", C.c_int32), ("reached", C.c_int64),
                ("residual", C.c_double), ("alg_bytes", C.c_double),
                ("kernel_ms", C.c_double), ("kernel_launches", C.c_int32),
                ("reserved", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class EpochTrace(C.Structure):
    _fields_ = [("count", C.c_int32), ("mode", C.c_int8 * 128),
                ("frontier", C.c_uint32 * 128), ("frontier_edges", C.c_uint64 * 128),
                ("scanned", C.c_uint64 * 128), ("us", C.c_double * 128)]


class SsspOpts(C.Structure):
    _fields_ = [("direction", C.c_int32), ("alpha", C.c_int32),
                ("max_epochs", C.c_int32)]


class BfsOpts(C.Structure):
    _fields_ = [("alpha", C.c_int32), ("beta", C.c_int32),
                ("direction", C.c_int32)]


class ScheduleBlock(C.Structure):
    _fields_ = [("device", C.c_int32), ("device_id", C.c_int32),
                ("start_segment", C.c_int32), ("end_segment", C.c_int32),
                ("kernel", C.c_int32), ("blocks", C.c_int32), ("threads", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_batch", C.c_int32), ("memory_budget", C.c_int64),
                ("min_batch_sssp", C.c_int32), ("min_batch_bfs", C.c_int32),
                ("max_wait_ms", C.c_int32), ("calibrate", C.c_int32),
                ("checksum", C.c_int32)]


class QueryInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("graph", C.c_int32),
                ("source", C.c_int32), ("batch_size", C.c_int32), ("slot", C.c_int32),
                ("num_nodes", C.c_int64), ("num_edges", C.c_int64), ("queue_ms", C.c_double),
                ("run_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("batches", C.c_int64), ("batched_queries", C.c_int64),
                ("uploads", C.c_int64), ("evictions", C.c_int64), ("spills", C.c_int64),
                ("bytes_in_use", C.c_int64), ("bytes_peak", C.c_int64)]


class PrOpts(C.Structure):
    _fields_ = [("damping", C.c_float), ("tolerance", C.c_double),
                ("max_iters", C.c_int32), ("rule", C.c_int32)]


_lib = None

_P = C.c_void_p
_SIGS = {
    "hc_last_error": ([], C.c_char_p),
    "hc_version": ([], C.c_int),
    "hc_init": ([C.c_int], C.c_int),
    "hc_get_device_info": ([C.POINTER(DeviceInfo)], C.c_int),
    "hc_device_synchronize": ([], C.c_int),
    "hc_mem_get_info": ([C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
    "hc_dmalloc": ([C.POINTER(_P), C.c_size_t], C.c_int),
    "hc_dfree": ([_P], C.c_int),
    "hc_hmalloc_pinned": ([C.POINTER(_P), C.c_size_t], C.c_int),
    "hc_hfree_pinned": ([_P], C.c_int),
    "hc_memcpy": ([_P, _P, C.c_size_t, C.c_int, _P], C.c_int),
    "hc_memset_d32": ([_P, C.c_int32, C.c_size_t, _P], C.c_int),
    "hc_stream_create": ([C.POINTER(_P)], C.c_int),
    "hc_stream_destroy": ([_P], C.c_int),
    "hc_stream_synchronize": ([_P], C.c_int),
    "hc_host_graph_generate": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64,
                                C.POINTER(HostGraphStruct)], C.c_int),
    "hc_host_graph_from_edges": ([_P, _P, _P, C.c_int64, C.c_int32, C.c_int,
                                  C.c_int, C.POINTER(HostGraphStruct)], C.c_int),
    "hc_host_graph_read": ([C.c_char_p, C.c_int, C.POINTER(HostGraphStruct)], C.c_int),
    "hc_host_graph_write": ([C.c_char_p, C.POINTER(HostGraphStruct), C.c_int], C.c_int),
    "hc_host_graph_free": ([C.POINTER(HostGraphStruct)], None),
    "hc_pick_sources": ([C.POINTER(HostGraphStruct), C.c_int64, C.c_int32, _P], C.c_int),
    "hc_equal_edge_ranges": ([C.POINTER(HostGraphStruct), C.c_int32, C.c_int32, _P], C.c_int),
    "hc_graph_upload": ([C.POINTER(HostGraphStruct), C.POINTER(_P)], C.c_int),
    "hc_graph_upload_range": ([C.POINTER(HostGraphStruct), C.c_int32, C.c_int32,
                               C.POINTER(_P)], C.c_int),
    "hc_graph_free": ([_P], C.c_int),
    "hc_graph_get_info": ([_P, C.POINTER(GraphInfo)], C.c_int),
    "hc_sssp_pull_epoch": ([_P, C.c_int, C.c_int32, C.c_int32, _P, _P, C.c_int,
                            C.c_int, _P], C.c_int),
    "hc_bfs_pull_epoch": ([_P, C.c_int, C.c_int32, C.c_int32, _P, _P, _P, _P,
                           C.c_int, C.c_int, _P], C.c_int),
    "hc_bfs_push_epoch": ([_P, _P, C.c_int32, _P, _P, _P, _P, _P], C.c_int),
    "hc_pr_pull_epoch": ([_P, C.c_int32, C.c_int32, C.c_float, _P, _P, _P, _P,
                          _P, _P, _P], C.c_int),
    "hc_sssp": ([_P, C.c_int32, _P, C.POINTER(SsspOpts), C.POINTER(RunStats), _P], C.c_int),
    "hc_sssp_pull": ([_P, C.c_int, _P, _P, C.c_int, C.c_int, C.POINTER(RunStats),
                      _P], C.c_int),
    "hc_bfs": ([_P, C.c_int32, _P, _P, C.POINTER(BfsOpts), C.POINTER(RunStats), _P], C.c_int),
    "hc_pagerank": ([_P, _P, C.POINTER(PrOpts), C.POINTER(RunStats), _P], C.c_int),
    "hc_pr_pull": ([_P, C.c_int, _P, _P, C.c_int, C.c_int, C.POINTER(PrOpts),
                    C.POINTER(RunStats), _P], C.c_int),
    "hc_get_epoch_trace": ([_P, C.c_int, C.POINTER(EpochTrace)], C.c_int),
    "hc_set_tuning": ([C.c_char_p, C.c_int64], C.c_int),
    "hc_get_tuning": ([C.c_char_p, C.POINTER(C.c_int64)], C.c_int),
    "hc_reset_tuning": ([], C.c_int),
    "hc_graph_set_tuning": ([_P, C.c_char_p, C.c_int64], C.c_int),
    "hc_graph_get_tuning": ([_P, C.c_char_p, C.POINTER(C.c_int64)], C.c_int),
    "hc_graph_reset_tuning": ([_P], C.c_int),
    "hc_sssp_multi": ([_P, _P, C.c_int32, _P, C.c_int32, C.POINTER(RunStats), _P], C.c_int),
    "hc_bfs_multi": ([_P, _P, C.c_int32, _P, C.c_int32, C.POINTER(RunStats), _P], C.c_int),
    "hc_graph_batch_upload": ([C.POINTER(HostGraphStruct), C.c_int32, C.POINTER(_P)], C.c_int),
    "hc_graph_batch_free": ([_P], C.c_int),
    "hc_graph_batch_info": ([_P, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                             C.POINTER(C.c_int64), C.POINTER(C.c_size_t)], C.c_int),
    "hc_graph_batch_node_offsets": ([_P, _P], C.c_int),
    "hc_bfs_many": ([_P, _P, _P, _P, C.c_int32, C.c_int32, C.POINTER(RunStats), _P], C.c_int),
    "hc_bfs_many_noalloc": ([_P, _P, _P, _P, C.c_int32, C.c_int32, C.POINTER(RunStats), _P],
                            C.c_int),
    "hc_graph_share": ([_P, C.POINTER(_P)], C.c_int),
    "hc_multi_extract": ([_P, C.c_int32, C.c_int32, C.c_int32, _P, _P], C.c_int),
    "hc_checksum_i32": ([_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                         C.POINTER(C.c_uint64), _P], C.c_int),
    "hc_checksum_multi": ([_P, C.c_int32, C.c_int32, C.c_int32, _P, _P], C.c_int),
    "hc_server_create": ([C.POINTER(ServerOpts), C.POINTER(_P)], C.c_int),
    "hc_server_destroy": ([_P], C.c_int),
    "hc_server_add_graph": ([_P, C.POINTER(HostGraphStruct), C.c_int32,
                             C.POINTER(C.c_int32)], C.c_int),
    "hc_server_set_graph_tuning": ([_P, C.c_int32, C.c_char_p, C.c_int64], C.c_int),
    "hc_server_graph_costs": ([_P, C.c_int32, _P, _P], C.c_int),
    "hc_server_submit": ([_P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64)], C.c_int),
    "hc_server_wait": ([_P, C.c_int64, C.POINTER(QueryInfo)], C.c_int),
    "hc_server_fetch": ([_P, C.c_int64, _P], C.c_int),
    "hc_server_release": ([_P, C.c_int64], C.c_int),
    "hc_server_checksum": ([_P, C.c_int64, C.POINTER(C.c_uint64)], C.c_int),
    "hc_server_get_stats": ([_P, C.POINTER(ServerStats)], C.c_int),
    "hc_server_graph_bytes": ([_P, C.POINTER(HostGraphStruct), C.c_int32,
                               C.POINTER(C.c_int64)], C.c_int),
    "hc_cpu_epoch_ms": ([C.POINTER(HostGraphStruct), C.c_int32, C.c_int32, C.c_int32,
                         C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "hc_kernel_from_string": ([C.c_char_p, C.POINTER(C.c_int32)], C.c_int),
    "hc_kernel_to_string": ([C.c_int32], C.c_char_p),
    "hc_kernel_device_type": ([C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "hc_device_from_string": ([C.c_char_p, C.POINTER(C.c_int32)], C.c_int),
    "hc_device_to_string": ([C.c_int32], C.c_char_p),
    "hc_device_type_of": ([C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "hc_schedule_parse": ([C.c_char_p, C.POINTER(_P)], C.c_int),
    "hc_schedule_load": ([C.c_char_p, C.POINTER(_P)], C.c_int),
    "hc_schedule_create": ([C.POINTER(ScheduleBlock), C.c_int32, C.c_int32,
                            C.POINTER(_P)], C.c_int),
    "hc_schedule_info": ([_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "hc_schedule_get_block": ([_P, C.c_int32, C.POINTER(ScheduleBlock)], C.c_int),
    "hc_schedule_to_text": ([_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)], C.c_int),
    "hc_schedule_free": ([_P], C.c_int),
    "hc_hetero_create": ([C.POINTER(HostGraphStruct), _P, C.POINTER(_P)], C.c_int),
    "hc_hetero_free": ([_P], C.c_int),
    "hc_sssp_heterogeneous": ([_P, _P, _P, C.POINTER(RunStats)], C.c_int),
    "hc_pr_heterogeneous": ([_P, C.POINTER(PrOpts), _P, C.POINTER(RunStats)], C.c_int),
    "hc_bfs_heterogeneous": ([_P, C.c_int32, C.c_int32, C.c_int32, _P, _P,
                              C.POINTER(RunStats)], C.c_int),
    "hc_cpu_sssp_pull_epoch": ([C.POINTER(HostGraphStruct), _P, C.c_int32, C.c_int32,
                                C.c_int, C.c_int, _P], None),
    "hc_cpu_pr_pull_epoch": ([C.POINTER(HostGraphStruct), C.c_float, _P, _P, _P, _P, _P,
                              C.c_int32, C.c_int32, C.c_int, C.c_int], C.c_double),
}

TUNING_KEYS = ("grid_per_cu", "sssp.alpha", "sssp.filter", "sssp.stage", "sssp.grid_per_cu",
               "sssp.block_threshold", "sssp.wave_threshold", "sssp.g16_threshold", "sssp.g4_threshold", "pull.stage", "bu.stage", "bu.remaining", "push.chunk", "push.stage", "bu.first", "sssp.dist16", "sssp.push_dedup", "compact.grid", "bu.stage_min", "loop.batch", "profile.kernel_events",
               "multi.graph", "multi.grid_per_cu", "pull.hub_threshold", "pull.hub_chunk",
               "pull.block_threshold", "pull.wave_threshold", "pull.g16_threshold",
               "pull.g4_threshold", "bu.wave_threshold", "bu.g16_threshold",
               "bu.g4_threshold", "multi.hub_threshold", "multi.hub_chunk",
               "multi.wave_threshold", "multi.g16_threshold", "pull.col_shift",
               "cb.min_slice", "cb.g4_slice", "cb.g16_slice", "cb.wave_slice", "cb.block_slice",
               "multi.block", "multi.delta", "multi.push_div", "sssp.delta", "many.block", "many.per_cu", "many.light_deg",
               "many.heavy_cap",
               "many.lds_nodes", "msbfs.on", "msbfs.alpha", "msbfs.col_shift")


def set_tuning(key: str, value: int):
    check(lib().hc_set_tuning(key.encode(), int(value)))


def get_tuning(key: str) -> int:
    v = C.c_int64()
    check(lib().hc_get_tuning(key.encode(), C.byref(v)))
    return v.value


def tuning() -> dict:
    return {k: get_tuning(k) for k in TUNING_KEYS}


def reset_tuning():
    check(lib().hc_reset_tuning())


def lib():
    """Loads libhcamd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != HC_OK:
        raise HcError(rc, lib().hc_last_error().decode(errors="replace"))
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_info() -> dict:
    d = DeviceInfo()
    check(lib().hc_get_device_info(C.byref(d)))
    return {"name": d.name.decode(), "arch": d.arch.decode(),
            "compute_units": d.compute_units, "wavefront_size": d.wavefront_size,
            "hbm_bytes": d.hbm_bytes, "lds_bytes_per_block": d.lds_bytes_per_block}


def mem_get_info() -> tuple[int, int]:
    """(free, total) device bytes of the current GPU."""
    f, t = C.c_size_t(), C.c_size_t()
    check(lib().hc_mem_get_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def synchronize():
    check(lib().hc_device_synchronize())


# ---------------------------------------------------------------------------
# Host graphs
# ---------------------------------------------------------------------------

class HostGraph:
    """CSR in host memory: index int64[N+1], neighbors/weights int32[m]."""

    def __init__(self, struct: HostGraphStruct, owned: bool, keep=None):
        self._s = struct
        self._owned = owned
        self._keep = keep

    # -- constructors --------------------------------------------------------
    @classmethod
    def generate(cls, kind: str, scale: int, degree: int, weighted: bool = True,
                 seed: int = 27491095) -> "HostGraph":
        if kind not in ("kron", "uniform"):
            raise ValueError("kind must be 'kron' or 'uniform'")
        s = HostGraphStruct()
        check(lib().hc_host_graph_generate(int(kind == "uniform"), scale, degree,
                                           int(weighted), seed, C.byref(s)))
        return cls(s, True)

    @classmethod
    def from_edges(cls, src, dst, num_nodes: int, weights=None,
                   symmetrize: bool = True, relabel: bool = False) -> "HostGraph":
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.int32)
        if src.shape != dst.shape or (w is not None and w.shape != src.shape):
            raise ValueError("edge arrays differ in length")
        s = HostGraphStruct()
        check(lib().hc_host_graph_from_edges(_ptr(src), _ptr(dst), _ptr(w), src.size,
                                             num_nodes, int(symmetrize), int(relabel),
                                             C.byref(s)))
        return cls(s, True)

    @classmethod
    def from_csr(cls, index, neighbors, weights=None) -> "HostGraph":
        """Wraps numpy CSR arrays (kept alive by the object, not copied)."""
        index = np.ascontiguousarray(index, dtype=np.int64)
        neighbors = np.ascontiguousarray(neighbors, dtype=np.int32)
        weights = None if weights is None else np.ascontiguousarray(weights, dtype=np.int32)
        s = HostGraphStruct()
        s.num_nodes = index.size - 1
        s.num_edges = neighbors.size
        s.index = index.ctypes.data_as(C.POINTER(C.c_int64))
        s.neighbors = neighbors.ctypes.data_as(C.POINTER(C.c_int32))
        s.weights = (weights.ctypes.data_as(C.POINTER(C.c_int32))
                     if weights is not None else C.POINTER(C.c_int32)())
        return cls(s, False, keep=(index, neighbors, weights))

    @classmethod
    def read(cls, path: str, weighted: bool) -> "HostGraph":
        s = HostGraphStruct()
        check(lib().hc_host_graph_read(path.encode(), int(weighted), C.byref(s)))
        return cls(s, True)

    def write(self, path: str, weighted: bool):
        check(lib().hc_host_graph_write(path.encode(), C.byref(self._s), int(weighted)))

    # -- accessors -----------------------------------------------------------
    @property
    def num_nodes(self) -> int:
        return self._s.num_nodes

    @property
    def num_edges(self) -> int:
        return self._s.num_edges

    @property
    def weighted(self) -> bool:
        return bool(self._s.weights)

    @property
    def index(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._s.index, shape=(self.num_nodes + 1,))

    @property
    def neighbors(self) -> np.ndarray:
        if self.num_edges == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self._s.neighbors, shape=(self.num_edges,))

    @property
    def weights(self):
        if not self.weighted:
            return None
        if self.num_edges == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self._s.weights, shape=(self.num_edges,))

    def degrees(self) -> np.ndarray:
        return np.diff(self.index)

    def pick_sources(self, count: int, seed: int = 37525) -> np.ndarray:
        out = np.zeros(count, np.int32)
        check(lib().hc_pick_sources(C.byref(self._s), seed, count, _ptr(out)))
        return out

    def equal_edge_ranges(self, segments: int, chunk: int = 1) -> np.ndarray:
        out = np.zeros(segments + 1, np.int32)
        check(lib().hc_equal_edge_ranges(C.byref(self._s), segments, chunk, _ptr(out)))
        return out

    def struct_ref(self):
        return C.byref(self._s)

    def free(self):
        if self._owned and self._s.index:
            lib().hc_host_graph_free(C.byref(self._s))
        self._keep = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Device memory and graphs
# ---------------------------------------------------------------------------

class DeviceArray:
    """Raw device allocation with numpy-typed host transfers."""

    def __init__(self, count: int, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        p = C.c_void_p()
        check(lib().hc_dmalloc(C.byref(p), max(1, self.count) * self.dtype.itemsize))
        self.ptr = p

    @classmethod
    def from_host(cls, arr) -> "DeviceArray":
        arr = np.ascontiguousarray(arr)
        d = cls(arr.size, arr.dtype)
        d.upload(arr)
        return d

    @property
    def nbytes(self):
        return self.count * self.dtype.itemsize

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.size == self.count
        check(lib().hc_memcpy(self.ptr, _ptr(arr), self.nbytes, 1, None))

    def get(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        check(lib().hc_memcpy(_ptr(out), self.ptr, self.nbytes, 2, None))
        return out

    def fill32(self, value: int):
        assert self.dtype.itemsize == 4
        v = int(np.array([value & 0xFFFFFFFF], np.uint32).view(np.int32)[0])
        check(lib().hc_memset_d32(self.ptr, v, self.count, None))

    def free(self):
        if self.ptr:
            lib().hc_dfree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceGraph:
    """CSR resident in HBM (whole graph or a row range)."""

    def __init__(self, host: HostGraph, start: int | None = None, end: int | None = None):
        h = C.c_void_p()
        if start is None and end is None:
            check(lib().hc_graph_upload(host.struct_ref(), C.byref(h)))
        else:
            start = 0 if start is None else start
            end = host.num_nodes if end is None else end
            check(lib().hc_graph_upload_range(host.struct_ref(), start, end, C.byref(h)))
        self.h = h
        self.num_nodes = host.num_nodes

    def share(self) -> "DeviceGraph":
        """A handle on the same device arrays with its own workspaces
        (hc_graph_share); this graph must outlive it."""
        h = C.c_void_p()
        check(lib().hc_graph_share(self.h, C.byref(h)))
        out = DeviceGraph.__new__(DeviceGraph)
        out.h = h
        out.num_nodes = self.num_nodes
        out._parent = self
        return out

    def set_tuning(self, key: str, value: int):
        """A knob of this graph and its shares only (hc_graph_set_tuning)."""
        check(lib().hc_graph_set_tuning(self.h, key.encode(), int(value)))

    def get_tuning(self, key: str) -> int:
        v = C.c_int64()
        check(lib().hc_graph_get_tuning(self.h, key.encode(), C.byref(v)))
        return v.value

    def reset_tuning(self):
        check(lib().hc_graph_reset_tuning(self.h))

    def info(self) -> dict:
        i = GraphInfo()
        check(lib().hc_graph_get_info(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    def free(self):
        if self.h:
            lib().hc_graph_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Full runs
# ---------------------------------------------------------------------------

@dataclass
class Result:
    values: np.ndarray
    stats: dict
    parents: np.ndarray | None = None


def sssp(g: DeviceGraph, source: int, direction: int = 0, alpha: int = 0,
         max_epochs: int = 0, out: DeviceArray | None = None,
         fetch: bool = True) -> Result:
    d = out if out is not None else DeviceArray(g.num_nodes, np.int32)
    st = RunStats()
    o = SsspOpts(direction, alpha, max_epochs)
    check(lib().hc_sssp(g.h, source, d.ptr, C.byref(o), C.byref(st), None))
    return Result(d.get() if fetch else None, st.as_dict())


def sssp_pull(g: DeviceGraph, init: np.ndarray, mapping: int = MAP_BUCKETED,
              blocks: int = 0, threads: int = 0) -> Result:
    di = DeviceArray.from_host(np.asarray(init, np.int32))
    d = DeviceArray(g.num_nodes, np.int32)
    st = RunStats()
    check(lib().hc_sssp_pull(g.h, mapping, di.ptr, d.ptr, blocks, threads,
                             C.byref(st), None))
    return Result(d.get(), st.as_dict())


def bfs(g: DeviceGraph, source: int, alpha: int = 15, beta: int = 18,
        direction: int = 0, want_parents: bool = True,
        depth_out: DeviceArray | None = None, fetch: bool = True) -> Result:
    dd = depth_out if depth_out is not None else DeviceArray(g.num_nodes, np.int32)
    dp = DeviceArray(g.num_nodes, np.int32) if want_parents else None
    st = RunStats()
    o = BfsOpts(alpha, beta, direction)
    check(lib().hc_bfs(g.h, source, dd.ptr, dp.ptr if dp else None, C.byref(o),
                       C.byref(st), None))
    return Result(dd.get() if fetch else None, st.as_dict(),
                  dp.get() if (dp is not None and fetch) else None)


def pagerank(g: DeviceGraph, damping: float = 0.85, tolerance: float = 1e-4,
             max_iters: int = 20, rule: int = 0) -> Result:
    """Jacobi PageRank; rule 0 stops on the FP64 L1 residual, rule 1 on
    pr_pull_cpu's per-vertex 5e-8 rule (stats['residual'] then counts the
    vertices that moved in the last iteration)."""
    d = DeviceArray(g.num_nodes, np.float32)
    st = RunStats()
    o = PrOpts(damping, tolerance, max_iters, rule)
    check(lib().hc_pagerank(g.h, d.ptr, C.byref(o), C.byref(st), None))
    return Result(d.get(), st.as_dict())


def pr_pull(g: DeviceGraph, mapping: int = MAP_BUCKETED, init=None, blocks: int = 0,
            threads: int = 0, damping: float = 0.85, tolerance: float = 1e-4,
            max_iters: int = 20, rule: int = 0) -> Result:
    """pr_pull_gpu's loop (hc_pr_pull): one epoch of the given mapping and
    launch shape per iteration, from `init` (host scores, or 1/N)."""
    d = DeviceArray(g.num_nodes, np.float32)
    di = DeviceArray.from_host(np.asarray(init, np.float32)) if init is not None else None
    st = RunStats()
    o = PrOpts(damping, tolerance, max_iters, rule)
    try:
        check(lib().hc_pr_pull(g.h, mapping, di.ptr if di else None, d.ptr, blocks, threads,
                               C.byref(o), C.byref(st), None))
        return Result(d.get(), st.as_dict())
    finally:
        d.free()
        if di:
            di.free()


# ---------------------------------------------------------------------------
# Epoch API
# ---------------------------------------------------------------------------

def sssp_pull_epoch(g: DeviceGraph, dist: DeviceArray, start: int, end: int,
                    mapping: int = MAP_BUCKETED, blocks: int = 0,
                    threads: int = 0) -> int:
    cnt = DeviceArray.from_host(np.zeros(1, np.int32))
    check(lib().hc_sssp_pull_epoch(g.h, mapping, start, end, dist.ptr, cnt.ptr,
                                   blocks, threads, None))
    return int(cnt.get()[0])


def bfs_pull_epoch(g: DeviceGraph, parents: DeviceArray, frontier: DeviceArray,
                   next_frontier: DeviceArray, start: int, end: int,
                   mapping: int = MAP_BUCKETED, blocks: int = 0,
                   threads: int = 0) -> int:
    cnt = DeviceArray.from_host(np.zeros(1, np.int32))
    check(lib().hc_bfs_pull_epoch(g.h, mapping, start, end, parents.ptr,
                                  frontier.ptr, next_frontier.ptr, cnt.ptr,
                                  blocks, threads, None))
    return int(cnt.get()[0])


def bfs_push_epoch(g: DeviceGraph, frontier: np.ndarray, parents: DeviceArray):
    f = DeviceArray.from_host(np.asarray(frontier, np.int32))
    nxt = DeviceArray(g.num_nodes, np.int32)
    nsz = DeviceArray.from_host(np.zeros(1, np.int32))
    ne = DeviceArray.from_host(np.zeros(1, np.int64))
    check(lib().hc_bfs_push_epoch(g.h, f.ptr, len(frontier), parents.ptr, nxt.ptr,
                                  nsz.ptr, ne.ptr, None))
    k = int(nsz.get()[0])
    return nxt.get()[:k], int(ne.get()[0])


def pr_pull_epoch(g: DeviceGraph, start: int, end: int, damping: float,
                  score_in: DeviceArray, contrib: DeviceArray, inv_degree: DeviceArray,
                  score_out: DeviceArray, contrib_out: DeviceArray) -> float:
    r = DeviceArray.from_host(np.zeros(1, np.float64))
    check(lib().hc_pr_pull_epoch(g.h, start, end, damping, score_in.ptr, contrib.ptr,
                                 inv_degree.ptr, score_out.ptr, contrib_out.ptr,
                                 r.ptr, None))
    return float(r.get()[0])


MULTI_WIDTH = 64


def sssp_multi(g: DeviceGraph, sources, out: DeviceArray | None = None,
               max_epochs: int = 0, fetch: bool = True) -> Result:
    """SSSP from up to 64 sources in one batch (hc_sssp_multi).  values is
    the [num_sources, N] distance matrix (row q from sources[q])."""
    src = np.ascontiguousarray(sources, np.int32)
    d = out if out is not None else DeviceArray(g.num_nodes * MULTI_WIDTH, np.int32)
    st = RunStats()
    check(lib().hc_sssp_multi(g.h, _ptr(src), src.size, d.ptr, max_epochs, C.byref(st),
                              None))
    vals = None
    if fetch:
        vals = d.get().reshape(g.num_nodes, MULTI_WIDTH)[:, :src.size].T.copy()
    return Result(vals, st.as_dict())


class GraphBatch:
    """Many host graphs resident in HBM as one batch (hc_graph_batch_upload)."""

    def __init__(self, graphs: list[HostGraph]):
        arr = (HostGraphStruct * len(graphs))(*[g._s for g in graphs])
        self.h = _P()
        check(lib().hc_graph_batch_upload(arr, len(graphs), C.byref(self.h)))
        ng, tn, te, nb = C.c_int32(), C.c_int64(), C.c_int64(), C.c_size_t()
        check(lib().hc_graph_batch_info(self.h, C.byref(ng), C.byref(tn), C.byref(te),
                                        C.byref(nb)))
        self.num_graphs, self.total_nodes, self.total_edges = ng.value, tn.value, te.value
        self.device_bytes = nb.value
        self.node_offsets = np.zeros(self.num_graphs + 1, np.int64)
        check(lib().hc_graph_batch_node_offsets(self.h, _ptr(self.node_offsets)))

    def bfs(self, sources, alpha: int = 15, beta: int = 18, want_parents: bool = True,
            depth: DeviceArray | None = None, parents: DeviceArray | None = None,
            fetch: bool = True, noalloc: bool = False):
        """One BFS per graph in one launch; returns (depths, parents, stats)
        with per-graph arrays split by node_offsets.  noalloc: run through
        hc_bfs_many_noalloc, which uses the library's fixed scratch instead
        of allocating bitmap slots."""
        src = np.ascontiguousarray(sources, np.int32)
        if src.size != self.num_graphs:
            raise ValueError(f"{src.size} sources for {self.num_graphs} graphs")
        dd = depth if depth is not None else DeviceArray(self.total_nodes, np.int32)
        dp = parents if parents is not None else (
            DeviceArray(self.total_nodes, np.int32) if want_parents else None)
        st = RunStats()
        fn = lib().hc_bfs_many_noalloc if noalloc else lib().hc_bfs_many
        check(fn(self.h, _ptr(src), dd.ptr, dp.ptr if dp else None, alpha, beta,
                 C.byref(st), None))
        if not fetch:
            return None, None, st.as_dict()
        o = self.node_offsets
        d = dd.get()
        p = dp.get() if dp is not None else None
        ds = [d[o[i]:o[i + 1]] for i in range(self.num_graphs)]
        ps = [p[o[i]:o[i + 1]] for i in range(self.num_graphs)] if p is not None else None
        return ds, ps, st.as_dict()

    def free(self):
        if self.h:
            lib().hc_graph_batch_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bfs_multi(g: DeviceGraph, sources, out: DeviceArray | None = None,
              max_epochs: int = 0, fetch: bool = True) -> Result:
    """BFS depths from up to 64 sources in one batch (hc_bfs_multi).  values
    is the [num_sources, N] depth matrix with -1 for unreachable vertices."""
    src = np.ascontiguousarray(sources, np.int32)
    d = out if out is not None else DeviceArray(g.num_nodes * MULTI_WIDTH, np.int32)
    st = RunStats()
    check(lib().hc_bfs_multi(g.h, _ptr(src), src.size, d.ptr, max_epochs, C.byref(st), None))
    vals = None
    if fetch:
        vals = d.get().reshape(g.num_nodes, MULTI_WIDTH)[:, :src.size].T.copy()
        vals[vals == INF_DIST] = -1
    return Result(vals, st.as_dict())


def checksum(d: DeviceArray, num_nodes: int, stride: int = 1, column: int = 0,
             unreached_minus1: bool = False) -> int:
    """hc_checksum_i32: order-independent 64-bit checksum of num_nodes int32
    values d[v * stride + column] (stride 1, or MULTI_WIDTH for a column of a
    batched result), computed on the device."""
    h = C.c_uint64()
    check(lib().hc_checksum_i32(d.ptr, int(num_nodes), int(stride), int(column),
                                int(unreached_minus1), C.byref(h), None))
    return h.value


def checksum_multi(d: DeviceArray, num_nodes: int, num_columns: int,
                   unreached_minus1: bool = False) -> np.ndarray:
    """hc_checksum_multi: the checksum() of columns 0..num_columns-1 of a
    batched [num_nodes][MULTI_WIDTH] result, in one pass."""
    out = np.zeros(int(num_columns), np.uint64)
    check(lib().hc_checksum_multi(d.ptr, int(num_nodes), int(num_columns),
                                  int(unreached_minus1), _ptr(out), None))
    return out


def epoch_trace(g: DeviceGraph, algorithm: str) -> list[dict]:
    """Per-epoch record of the last sssp/bfs run on g."""
    t = EpochTrace()
    check(lib().hc_get_epoch_trace(g.h, 0 if algorithm == "sssp" else 1, C.byref(t)))
    names = {0: "push", 1: "pull", 2: "promote"}
    return [{"mode": names.get(t.mode[i], str(t.mode[i])), "frontier": t.frontier[i],
             "frontier_edges": t.frontier_edges[i], "scanned": t.scanned[i],
             "us": round(t.us[i], 2)}
            for i in range(t.count)]


def gteps(num_edges_directed: int, ms: float) -> float:
    """Reference GTEPS: m / 2 / t (src/benchmarks/gpu_benchmark.cuh:590)."""
    return num_edges_directed / 2.0 / (ms / 1e3) / 1e9


# ---------------------------------------------------------------------------
# Segment schedules and host+GPU co-execution
# ---------------------------------------------------------------------------

KERNELS = ("SSSP CPU one-to-one", "SSSP GPU one-to-one", "SSSP GPU warp-red",
           "SSSP GPU block-red", "SSSP GPU bucketed", "PR CPU one-to-one",
           "PR GPU one-to-one", "PR GPU warp-red", "PR GPU block-red", "PR GPU bucketed",
           "BFS CPU pull", "BFS GPU one-to-one", "BFS GPU warp", "BFS GPU bucketed",
           "BFS CPU push by node", "BFS CPU push by edge")
DEVICES = ("Intel i7-9700K", "Intel Xeon E5-2686", "NVIDIA Quadro RTX 4000",
           "NVIDIA Tesla K80", "NVIDIA Tesla M60", "AMD Instinct MI355X", "Host CPU")
DEV_CPU, DEV_GPU = 0, 1


def kernel_id(name: str) -> int:
    k = C.c_int32()
    check(lib().hc_kernel_from_string(name.encode(), C.byref(k)))
    return k.value


def device_id(name: str) -> int:
    d = C.c_int32()
    check(lib().hc_device_from_string(name.encode(), C.byref(d)))
    return d.value


class Schedule:
    """A segment schedule (.skd), see hc_schedule_parse."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def parse(cls, text: str) -> "Schedule":
        h = _P()
        check(lib().hc_schedule_parse(text.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path: str) -> "Schedule":
        h = _P()
        check(lib().hc_schedule_load(os.fsencode(path), C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, blocks: list[dict], num_segments: int) -> "Schedule":
        """blocks: dicts with device (name), start/end segment (inclusive),
        kernel (name), optional device_id, blocks, threads."""
        arr = (ScheduleBlock * len(blocks))()
        for i, b in enumerate(blocks):
            arr[i] = ScheduleBlock(device_id(b["device"]), b.get("device_id", 0),
                                   b["start"], b["end"], kernel_id(b["kernel"]),
                                   b.get("blocks", 0), b.get("threads", 0))
        h = _P()
        check(lib().hc_schedule_create(arr, len(blocks), num_segments, C.byref(h)))
        return cls(h)

    @property
    def num_segments(self) -> int:
        nb, ns = C.c_int32(), C.c_int32()
        check(lib().hc_schedule_info(self.h, C.byref(nb), C.byref(ns)))
        return ns.value

    @property
    def blocks(self) -> list[dict]:
        nb, ns = C.c_int32(), C.c_int32()
        check(lib().hc_schedule_info(self.h, C.byref(nb), C.byref(ns)))
        out = []
        for i in range(nb.value):
            b = ScheduleBlock()
            check(lib().hc_schedule_get_block(self.h, i, C.byref(b)))
            out.append({"device": DEVICES[b.device], "device_id": b.device_id,
                        "start": b.start_segment, "end": b.end_segment,
                        "kernel": KERNELS[b.kernel], "blocks": b.blocks,
                        "threads": b.threads})
        return out

    def text(self) -> str:
        need = C.c_size_t()
        check(lib().hc_schedule_to_text(self.h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        check(lib().hc_schedule_to_text(self.h, buf, need.value, C.byref(need)))
        return buf.value.decode()

    def free(self):
        if self.h:
            lib().hc_schedule_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Hetero:
    """A schedule bound to a host graph (hc_hetero_create).  The host graph
    must outlive this object."""

    def __init__(self, g: HostGraph, schedule: Schedule):
        self._g = g
        self.h = _P()
        check(lib().hc_hetero_create(g.struct_ref(), schedule.h, C.byref(self.h)))

    def sssp(self, source: int | None = None, init: np.ndarray | None = None) -> Result:
        n = self._g.num_nodes
        if init is None:
            init = np.full(n, INF_DIST, np.int32)
            init[source] = 0
        init = np.ascontiguousarray(init, np.int32)
        out = np.empty(n, np.int32)
        st = RunStats()
        check(lib().hc_sssp_heterogeneous(self.h, _ptr(init), _ptr(out), C.byref(st)))
        return Result(out, st.as_dict())

    def pagerank(self, damping: float = 0.85, tolerance: float = 1e-4,
                 max_iters: int = 20) -> Result:
        out = np.empty(self._g.num_nodes, np.float32)
        st = RunStats()
        o = PrOpts(damping, tolerance, max_iters)
        check(lib().hc_pr_heterogeneous(self.h, C.byref(o), _ptr(out), C.byref(st)))
        return Result(out, st.as_dict())

    def bfs(self, source: int, alpha: int = 15, beta: int = 18) -> Result:
        n = self._g.num_nodes
        par = np.empty(n, np.int32)
        dep = np.empty(n, np.int32)
        st = RunStats()
        check(lib().hc_bfs_heterogeneous(self.h, source, alpha, beta, _ptr(par),
                                         _ptr(dep), C.byref(st)))
        return Result(dep, st.as_dict(), par)

    def free(self):
        if self.h:
            lib().hc_hetero_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Serving loop
# ---------------------------------------------------------------------------

QUERY_BFS, QUERY_SSSP = 0, 1


class Server:
    """Resident/dynamic graphs and continuously batched BFS/SSSP queries
    (hc_server_*).  Host graphs added here are kept alive by the object."""

    def __init__(self, slots: int = 4, max_batch: int = 64, memory_budget: int = 0,
                 min_batch_sssp: int = 0, min_batch_bfs: int = 0, max_wait_ms: int = 0,
                 calibrate: int = 0, checksum: bool = False):
        """min_batch_*: queries of one graph and kind that must wait before
        they run as one batched sweep (0: from the costs measured when a
        resident graph is added, else 8 and 64); max_wait_ms: how long
        fast-arriving queries may wait to gather a batch (0: the default,
        500; -1: never); calibrate -1: measure no costs; checksum: each
        slot sums every result of its sweep on the device (checksum())."""
        self.h = _P()
        o = ServerOpts(slots, max_batch, memory_budget, min_batch_sssp, min_batch_bfs,
                       max_wait_ms, calibrate, int(checksum))
        check(lib().hc_server_create(C.byref(o), C.byref(self.h)))
        self._graphs = []

    def add_graph(self, g: HostGraph, resident: bool = True) -> int:
        gid = C.c_int32()
        check(lib().hc_server_add_graph(self.h, g.struct_ref(), int(resident), C.byref(gid)))
        self._graphs.append(g)
        return gid.value

    def graph_costs(self, graph: int) -> dict:
        """Device ms of a single query and of a 64-wide batch per kind,
        measured when the graph was added (hc_server_graph_costs)."""
        sm = (C.c_double * 2)()
        bm = (C.c_double * 2)()
        check(lib().hc_server_graph_costs(self.h, graph, sm, bm))
        return {"bfs": {"single_ms": sm[0], "batch_ms": bm[0]},
                "sssp": {"single_ms": sm[1], "batch_ms": bm[1]}}

    def set_graph_tuning(self, graph: int, key: str, value: int):
        """A knob for queries on one graph (hc_server_set_graph_tuning)."""
        check(lib().hc_server_set_graph_tuning(self.h, graph, key.encode(), int(value)))

    def submit(self, graph: int, kind: int, source: int) -> int:
        t = C.c_int64()
        check(lib().hc_server_submit(self.h, graph, kind, int(source), C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> dict:
        """Query info; a failed query has status != 0 and "error" set."""
        i = QueryInfo()
        check(lib().hc_server_wait(self.h, ticket, C.byref(i)))
        d = {f: getattr(i, f) for f, _ in i._fields_}
        if d["status"] != HC_OK:
            d["error"] = lib().hc_last_error().decode(errors="replace")
        return d

    def fetch(self, ticket: int) -> np.ndarray:
        info = self.wait(ticket)
        out = np.empty(info["num_nodes"], np.int32)
        check(lib().hc_server_fetch(self.h, ticket, _ptr(out)))
        return out

    def checksum(self, ticket: int) -> int:
        """hc_checksum_i32 of the query's result as fetch() returns it,
        computed where the result lies (hc_server_checksum)."""
        h = C.c_uint64()
        check(lib().hc_server_checksum(self.h, ticket, C.byref(h)))
        return h.value

    def release(self, ticket: int):
        check(lib().hc_server_release(self.h, ticket))

    def graph_bytes(self, g: HostGraph, resident: bool) -> int:
        """Bytes the memory budget charges for g (hc_server_graph_bytes)."""
        b = C.c_int64()
        check(lib().hc_server_graph_bytes(self.h, g.struct_ref(), int(resident), C.byref(b)))
        return b.value

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().hc_server_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if self.h:
            lib().hc_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
