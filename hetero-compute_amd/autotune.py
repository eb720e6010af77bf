"""Autotuner: measured choices of degree-bucket mapping, launch shape and
host/GPU segment assignment.

The reference chooses kernels and schedules offline: per-segment
microbenchmarks write YAML profiles (src/benchmarks/benchmark.cuh:87-309),
a Python scheduler assigns segments to devices
(util/scheduler/scheduler.py:237-278, MostGainScheduler) and a code
generator emits one driver per schedule (kernelgen/sssp_hetero.py).  Here
the measurements drive the library directly:

* ``tune_knobs`` searches the tuning knobs (workgroup size and workgroups
  per CU, degree-class thresholds, LDS staging, column blocking, SSSP
  push/pull ratio, batched-sweep classes, the many-graph kernel's launch
  shape and per-level thresholds) by coordinate descent on measured device
  time, checking that every candidate returns bit-identical results, and
  applies the best, process-wide or to the one graph (``per_graph``);
* ``profile_segments`` times every equal-edge segment on the host cores and
  with every GPU epoch mapping and launch shape;
* ``schedule_from_profile`` picks each segment's fastest GPU kernel and moves
  segments to the host cores while that lowers the per-epoch makespan
  (GPU time + PCIe exchange vs host time), then merges neighbouring segments
  into blocks: the result is an ``hc`` Schedule (``.skd`` text) that
  ``Hetero`` runs;
* results are cached in a JSON file keyed by device, algorithm and graph
  signature; ``apply_cached`` sets a graph's cached knobs on its handle (or
  on a server's graph), so callers running several graphs each get their
  own.

Run as a script to tune a generated graph:
    python hetero-compute_amd/autotune.py --kind kron --scale 20 --algorithm sssp
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

if __package__:
    hc = sys.modules[__package__]
else:  # run as a script
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import importlib

    hc = importlib.import_module("hetero-compute_amd")

# Candidate values per knob, searched one knob at a time in this order.
KNOB_SPACE = {
    "sssp": {
        "sssp.alpha": [2, 4, 8, 16],
        "sssp.filter": [0, 1],
        "sssp.dist16": [0, 1],
        "sssp.push_dedup": [0, 1],
        "sssp.stage": [0, 2048, 8192],
        "sssp.block_threshold": [1024, 4096, 16384],
        "sssp.wave_threshold": [48, 96, 192],
        "sssp.g16_threshold": [12, 24, 48],
        "sssp.g4_threshold": [5, 8, 16],
        "pull.hub_threshold": [16384, 65536],
        "sssp.grid_per_cu": [2, 4, 8, 16],
        "pull.col_shift": [-1, 0, 20, 21],
        "cb.min_slice": [2, 4, 8],
        "cb.g16_slice": [8, 12, 24],
        "sssp.delta": [-1, 0, 8, 32],
    },
    "bfs": {
        "grid_per_cu": [2, 4, 8, 16],
        "bu.stage": [0, 2048, 8192, 15360],
        "bu.wave_threshold": [1024, 4096],
        "bu.g16_threshold": [128, 256, 512],
        "bu.g4_threshold": [8, 16, 32, 64],
        "bu.remaining": [0, 1, 4],
        "bu.stage_min": [0, 256, 1024, 4096, 1 << 20],
        "bu.first": [0, 1, 2],
        "push.chunk": [256, 512, 1024],
        "compact.grid": [128, 512, 2048],
    },
    "pr": {
        "grid_per_cu": [2, 4, 8, 16],
        "pull.stage": [0, 8192, 15360],
        "pull.wave_threshold": [32, 96, 192],
        "pull.g16_threshold": [8, 24, 48],
        "pull.hub_threshold": [4096, 16384, 65536],
    },
    "multi": {
        "multi.block": [128, 256, 512, 1024],
        "multi.grid_per_cu": [4, 8, 16],
        "multi.wave_threshold": [8, 16, 32, 64],
        "multi.g16_threshold": [2, 4, 8],
        "multi.hub_threshold": [2048, 4096, 16384],
        "multi.hub_chunk": [1024, 2048, 8192],
        "multi.push_div": [0, 4, 16, 64],
        "multi.delta": [-1, 0, 48, 96, 192],
    },
    # The 64-source bit-parallel BFS.
    "bfs_multi": {
        "msbfs.alpha": [8, 16, 32, 64, 128],
        "msbfs.col_shift": [-1, 0, 18, 19, 20, 21],
    },
    # The many-graph BFS (process-wide knobs; `dg` is a GraphBatch).
    "many": {
        "many.block": [256, 512, 1024],
        "many.per_cu": [0, 1, 2, 4],
        "many.light_deg": [8, 16, 32, 64, 128],
        "many.heavy_cap": [128, 256, 512, 1024],
        "many.lds_nodes": [0, 1 << 16, 1 << 17],
    },
}

GPU_KERNELS = {
    "sssp": ["SSSP GPU one-to-one", "SSSP GPU warp-red", "SSSP GPU block-red",
             "SSSP GPU bucketed"],
    "bfs": ["BFS GPU one-to-one", "BFS GPU warp", "BFS GPU bucketed"],
    "pr": ["PR GPU bucketed"],
}
CPU_KERNEL = {"sssp": "SSSP CPU one-to-one", "bfs": "BFS CPU pull", "pr": "PR CPU one-to-one"}
SHAPES = [(0, 0), (256, 256), (1024, 256), (512, 512), (256, 1024)]
MAPPING = {"one-to-one": 0, "warp-red": 1, "warp": 1, "block-red": 2, "bucketed": 3}


def signature(g) -> str:
    """Graph key for the cache: size and a hash of the degree sequence (of
    every graph, for a list of graphs: a batch)."""
    if isinstance(g, (list, tuple)):
        h = hashlib.sha1()
        for x in g:
            h.update(np.diff(x.index).astype(np.int64).tobytes())
        return (f"batch{len(g)}-n{sum(x.num_nodes for x in g)}-"
                f"m{sum(x.num_edges for x in g)}-{h.hexdigest()[:16]}")
    deg = np.diff(g.index).astype(np.int64)
    h = hashlib.sha1(deg.tobytes()).hexdigest()[:16]
    return f"n{g.num_nodes}-m{g.num_edges}-{h}"


# ---------------------------------------------------------------------------
# Knob search
# ---------------------------------------------------------------------------

def _run(dg, algorithm, sources):
    """Mean device ms over sources and the concatenated results."""
    ms, outs = [], []
    for s in sources:
        if algorithm == "sssp":
            r = hc.sssp(dg, int(s))
        elif algorithm == "bfs":
            r = hc.bfs(dg, int(s), want_parents=False)
        elif algorithm == "pr":
            r = hc.pagerank(dg)
        elif algorithm == "many":
            d, _, st = dg.bfs(sources, want_parents=False)
            ms.append(st["ms"])
            outs.extend(d)
            break
        elif algorithm == "bfs_multi":
            r = hc.bfs_multi(dg, sources)
            ms.append(r.stats["ms"])
            outs.append(r.values)
            break
        else:
            r = hc.sssp_multi(dg, sources)
            ms.append(r.stats["ms"])
            outs.append(r.values)
            break
        ms.append(r.stats["ms"])
        outs.append(r.values)
    return float(np.mean(ms)), outs


def _same(a, b, algorithm):
    if algorithm == "pr":  # summation order may change with the mapping
        return all(np.abs(x.astype(np.float64) - y).sum() <= 1e-6 for x, y in zip(a, b))
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def tune_knobs(dg, algorithm: str, sources, space: dict | None = None, repeats: int = 2,
               log=None, per_graph: bool = False) -> dict:
    """Coordinate descent over the knobs of `algorithm` ("sssp", "bfs", "pr",
    "multi" for the batched SSSP sweep, "bfs_multi" for the 64-source
    bit-parallel BFS, or "many" for the many-graph BFS, where
    `dg` is a GraphBatch and `sources` one source per graph).  Every
    candidate's results must match the default configuration's exactly
    (PageRank: within 1e-6 L1); the best setting found is applied and
    returned with its timings.  per_graph: knobs are set on `dg` only
    (DeviceGraph.set_tuning), not process-wide."""
    space = space or KNOB_SPACE[algorithm]
    if per_graph and algorithm == "many":
        raise ValueError("the many-graph knobs are process-wide")
    setk = dg.set_tuning if per_graph else hc.set_tuning
    getk = dg.get_tuning if per_graph else hc.get_tuning
    _run(dg, algorithm, sources)  # warm plans and workspaces
    base_ms, base_out = min((_run(dg, algorithm, sources) for _ in range(repeats)),
                            key=lambda t: t[0])
    best = {k: getk(k) for k in space}
    base_knobs = dict(best)
    best_ms = base_ms
    trials = []
    for key, values in space.items():
        for v in values:
            if v == best[key]:
                continue
            setk(key, v)
            ms, out = min((_run(dg, algorithm, sources) for _ in range(repeats)),
                          key=lambda t: t[0])
            ok = _same(out, base_out, algorithm)
            trials.append({"knob": key, "value": v, "ms": round(ms, 4), "exact": ok})
            if log:
                log(f"  {key}={v}: {ms:.3f} ms{'' if ok else ' (MISMATCH)'}")
            if not ok:
                raise RuntimeError(f"{key}={v} changed the results of {algorithm}")
            if ms < best_ms:
                best_ms, best[key] = ms, v
        setk(key, best[key])
    # The winner timed again, with every chosen knob in place; when that is
    # not faster than the starting configuration (single-knob gains that do
    # not add up, or timing noise), the starting knobs are kept.
    final_ms, final_out = min((_run(dg, algorithm, sources) for _ in range(repeats)),
                              key=lambda t: t[0])
    if not _same(final_out, base_out, algorithm):
        raise RuntimeError(f"the chosen knobs changed the results of {algorithm}")
    if final_ms > base_ms:
        for key in space:
            best[key] = base_knobs[key]
            setk(key, best[key])
        final_ms = base_ms
    return {"algorithm": algorithm, "default_ms": round(base_ms, 4),
            "tuned_ms": round(final_ms, 4), "knobs": best, "trials": trials}


def apply_knobs(knobs: dict, target=None):
    """Sets knobs process-wide, on a DeviceGraph, or on a server's graph
    (target = (server, graph_id))."""
    for k, v in knobs.items():
        if target is None:
            hc.set_tuning(k, int(v))
        elif isinstance(target, tuple):
            target[0].set_graph_tuning(target[1], k, int(v))
        else:
            target.set_tuning(k, int(v))


def apply_cached(path: str, g, algorithm: str, target=None) -> dict:
    """Applies the knobs cached for graph `g` (a HostGraph, or a list of
    them for "many") and `algorithm` to `target` (see apply_knobs); returns
    them ({} when the cache holds none for this device and graph)."""
    entry = load(path, cache_key(g, algorithm))
    knobs = dict(entry["knobs"]) if entry else {}
    if knobs:
        apply_knobs(knobs, target)
    return knobs


# ---------------------------------------------------------------------------
# Segment profiling and scheduling
# ---------------------------------------------------------------------------

def _time_gpu_epoch(dg, algorithm, start, end, mapping, blocks, threads, state, aux, repeats):
    best = float("inf")
    for _ in range(repeats):
        hc.synchronize()
        t0 = time.perf_counter()
        if algorithm == "sssp":
            hc.sssp_pull_epoch(dg, state, start, end, mapping=mapping, blocks=blocks,
                               threads=threads)
        elif algorithm == "bfs":
            hc.bfs_pull_epoch(dg, state, aux[0], aux[1], start, end, mapping=mapping,
                              blocks=blocks, threads=threads)
        else:
            hc.pr_pull_epoch(dg, start, end, 0.85, *aux)
        hc.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def profile_segments(g, algorithm: str, num_segments: int = 24, shapes=None,
                     repeats: int = 3) -> dict:
    """Per-segment epoch times: host cores (all threads) and every GPU
    kernel/launch shape, on range handles of the segments."""
    shapes = shapes or SHAPES
    bounds = g.equal_edge_ranges(num_segments)
    n = g.num_nodes
    alg_id = {"sssp": 0, "bfs": 1, "pr": 2}[algorithm]
    import ctypes as C
    segs = []
    with hc.DeviceGraph(g) as dg:
        if algorithm == "sssp":
            state = hc.DeviceArray.from_host(np.ones(n, np.int32))
            aux = None
        elif algorithm == "bfs":
            state = hc.DeviceArray.from_host(np.full(n, -1, np.int32))
            words = (n + 31) // 32
            aux = (hc.DeviceArray.from_host(np.zeros(words, np.uint32)),
                   hc.DeviceArray.from_host(np.zeros(words, np.uint32)))
        else:
            deg = np.diff(g.index)
            inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
            s0 = np.full(n, 1.0 / n, np.float32)
            aux = (hc.DeviceArray.from_host(s0),
                   hc.DeviceArray.from_host((s0 * inv).astype(np.float32)),
                   hc.DeviceArray.from_host(inv),
                   hc.DeviceArray(n, np.float32), hc.DeviceArray(n, np.float32))
            state = None
        for k in range(num_segments):
            a, b = int(bounds[k]), int(bounds[k + 1])
            ms = C.c_double()
            hc.check(hc.lib().hc_cpu_epoch_ms(g.struct_ref(), alg_id, a, b, repeats,
                                              C.byref(ms)))
            gpu = {}
            for name in GPU_KERNELS[algorithm]:
                mapping = MAPPING[name.split()[-1]]
                cand = [(0, 0)] if mapping == 3 or algorithm == "pr" else shapes
                for blocks, threads in cand:
                    if mapping != 0 and threads % 64:
                        continue
                    t = _time_gpu_epoch(dg, algorithm, a, b, mapping, blocks, threads, state,
                                        aux, repeats)
                    if name not in gpu or t < gpu[name][0]:
                        gpu[name] = (t, blocks, threads)
            segs.append({"start": a, "end": b, "edges": int(g.index[b] - g.index[a]),
                         "cpu_ms": ms.value, "gpu": gpu})
    return {"algorithm": algorithm, "num_nodes": n, "segments": segs}


def profile_tree(g, algorithm: str, depth: int, shapes=None, repeats: int = 3) -> dict:
    """Tree microbenchmark (src/benchmarks/benchmark.cuh:258-291): layer i
    splits the vertices into 2^i equal-edge segments and profiles each one on
    the host cores and every GPU kernel, for i = 0..depth."""
    if depth < 0:
        raise ValueError("depth must be >= 0")
    layers = [profile_segments(g, algorithm, 1 << i, shapes, repeats) for i in range(depth + 1)]
    return {"algorithm": algorithm, "num_nodes": g.num_nodes, "layers": layers}


def schedule_from_profile(profile: dict, pcie_gbs: float = 25.0):
    """Greedy host/GPU split of the profiled segments.  The GPU side costs
    the sum of its segments' best kernel times plus, once any segment is on
    the host, the per-epoch exchange of 4 bytes per vertex in each direction
    over PCIe; the host side costs the sum of its segments' times.  Segments
    move to the host, best GPU-to-CPU time ratio first, while the makespan
    drops.  Returns (blocks, makespan_ms, plan) with blocks ready for
    Schedule.create."""
    segs = profile["segments"]
    alg = profile["algorithm"]
    n = profile["num_nodes"]
    best_gpu = []
    for s in segs:
        name, (t, blocks, threads) = min(s["gpu"].items(), key=lambda kv: kv[1][0])
        best_gpu.append((name, t, blocks, threads))
    on_cpu = [False] * len(segs)

    def makespan(flags):
        gpu_t = sum(best_gpu[i][1] for i, f in enumerate(flags) if not f)
        cpu_t = sum(segs[i]["cpu_ms"] for i, f in enumerate(flags) if f)
        if any(flags):
            gpu_t += 2.0 * 4.0 * n / (pcie_gbs * 1e9) * 1e3
        return max(gpu_t, cpu_t)

    cur = makespan(on_cpu)
    order = sorted(range(len(segs)), key=lambda i: -best_gpu[i][1] / max(segs[i]["cpu_ms"], 1e-9))
    for i in order:
        trial = list(on_cpu)
        trial[i] = True
        m = makespan(trial)
        if m < cur:
            cur, on_cpu = m, trial
    # Merge neighbouring segments with the same device and kernel.
    blocks = []
    for i, s in enumerate(segs):
        if on_cpu[i]:
            dev, kern, b, t = "Host CPU", CPU_KERNEL[alg], 0, 0
        else:
            dev, kern, b, t = "AMD Instinct MI355X", best_gpu[i][0], best_gpu[i][2], best_gpu[i][3]
        if blocks and blocks[-1]["device"] == dev and blocks[-1]["kernel"] == kern and \
                blocks[-1]["blocks"] == b and blocks[-1]["threads"] == t and \
                blocks[-1]["end"] == i - 1:
            blocks[-1]["end"] = i
        else:
            blocks.append({"device": dev, "start": i, "end": i, "kernel": kern, "blocks": b,
                           "threads": t})
    plan = {"cpu_segments": [i for i, f in enumerate(on_cpu) if f],
            "gpu_kernels": [x[0] for x in best_gpu]}
    return blocks, cur, plan


def _segment_rows(g, segments):
    deg = np.diff(g.index)
    rows = []
    for sgm in segments:
        d = deg[sgm["start"]:sgm["end"]]
        rows.append({"start": sgm["start"], "end": sgm["end"],
                     "avg_deg": float(d.mean()) if d.size else 0.0,
                     "min_deg": int(d.min()) if d.size else 0,
                     "max_deg": int(d.max()) if d.size else 0,
                     "num_edges": sgm["edges"]})
    return rows


def _segment_yaml(rows, millis, indent):
    pad = " " * indent
    out = []
    for r, ms in zip(rows, millis):
        gteps = r["num_edges"] / 2 / (ms / 1e3) / 1e9 if ms > 0 else 0.0
        out += [f"{pad}- start: {r['start']}", f"{pad}  end: {r['end']}",
                f"{pad}  avg_deg: {r['avg_deg']:.6g}", f"{pad}  min_deg: {r['min_deg']}",
                f"{pad}  max_deg: {r['max_deg']}", f"{pad}  num_edges: {r['num_edges']}",
                f"{pad}  millis: {ms:.6g}", f"{pad}  gteps: {gteps:.6g}"]
    return out


def _kernel_series(segments, alg, cpu_name, gpu_name):
    """(file tag, device, kernel line, per-segment ms) for the host kernel and
    each GPU kernel profiled on every segment (its most common best shape)."""
    series = [(f"cpu_{CPU_KERNEL[alg].split()[-1]}", cpu_name, CPU_KERNEL[alg],
               [sg["cpu_ms"] for sg in segments])]
    for name in GPU_KERNELS[alg]:
        if not all(name in sg["gpu"] for sg in segments):
            continue
        shapes = [sg["gpu"][name][1:] for sg in segments]
        b, t = max(set(shapes), key=shapes.count)
        series.append((f"gpu_{name.split()[-1]}", gpu_name, f"{name} {b} {t}",
                       [sg["gpu"][name][0] for sg in segments]))
    return series


def write_yaml_tree(tree: dict, g, outdir: str, cpu_name: str = "Host CPU",
                    gpu_name: str = "AMD Instinct MI355X") -> list[str]:
    """Writes a tree profile in the layout of the reference's tree_res_t
    writer (src/benchmarks/benchmark.cuh:213-240): per device and kernel,
    `config: {device, kernel}` and `results: - depth: i, segments: [...]`."""
    os.makedirs(outdir, exist_ok=True)
    alg = tree["algorithm"]
    per_file = {}
    for depth, layer in enumerate(tree["layers"]):
        rows = _segment_rows(g, layer["segments"])
        for tag, dev, kname, millis in _kernel_series(layer["segments"], alg, cpu_name, gpu_name):
            entry = per_file.setdefault(tag, {"device": dev, "kernel": kname, "layers": []})
            entry["layers"].append((depth, rows, millis))
    out = []
    for tag, e in per_file.items():
        if len(e["layers"]) != len(tree["layers"]):
            continue  # a kernel missing from some layer has no complete tree
        lines = ["config:", f"  device: {e['device']}", f"  kernel: {e['kernel']}", "results:"]
        for depth, rows, millis in e["layers"]:
            lines += [f"  - depth: {depth}", "    segments:"] + _segment_yaml(rows, millis, 6)
        path = os.path.join(outdir, f"{alg}_{tag}_tree.yaml")
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
        out.append(path)
    return out


def write_yaml_profiles(profile: dict, g, outdir: str, cpu_name: str = "Host CPU",
                        gpu_name: str = "AMD Instinct MI355X") -> list[str]:
    """Writes the segment profile as the reference's per-kernel YAML files
    (layer results, src/benchmarks/benchmark.cuh:184-207): one file per
    device and kernel with `config: {device, kernel}` and `results: -
    segments: [...]`, the input of util/scheduler/sched.py.  Launch shapes
    are part of the GPU kernel name, as the reference's profiles carry them."""
    os.makedirs(outdir, exist_ok=True)
    alg = profile["algorithm"]
    rows = _segment_rows(g, profile["segments"])
    out = []
    for tag, dev, kname, millis in _kernel_series(profile["segments"], alg, cpu_name, gpu_name):
        lines = ["config:", f"  device: {dev}", f"  kernel: {kname}", "results:",
                 "  - segments:"] + _segment_yaml(rows, millis, 6)
        path = os.path.join(outdir, f"{alg}_{tag}.yaml")
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
        out.append(path)
    return out

def tune_schedule(g, algorithm: str, num_segments: int = 24, pcie_gbs: float = 25.0):
    prof = profile_segments(g, algorithm, num_segments)
    blocks, makespan, plan = schedule_from_profile(prof, pcie_gbs)
    return hc.Schedule.create(blocks, num_segments), {"makespan_ms": makespan, **plan,
                                                      "profile": prof}


# ---------------------------------------------------------------------------
# Cache
# ---------------------------------------------------------------------------

def cache_key(g, algorithm: str) -> str:
    try:
        dev = hc.device_info()["name"]
    except Exception:
        dev = "unknown"
    return f"{dev}|{algorithm}|{signature(g)}"


def save(path: str, key: str, entry: dict):
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = entry
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def load(path: str, key: str) -> dict | None:
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


def many_graphs(count: int, seed: int = 4242):
    """The many-graph workload: `count` Kronecker graphs of scale U{10..16}
    and edge factor U{8..32}, graph i from seed + 1 + i."""
    rng = np.random.RandomState(seed)
    return [hc.HostGraph.generate("kron", int(rng.randint(10, 17)), int(rng.randint(8, 33)),
                                  weighted=False, seed=seed + 1 + i) for i in range(count)]


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--kind", default="kron")
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--degree", type=int, default=16)
    ap.add_argument("--algorithm", default="sssp",
                    choices=["sssp", "bfs", "pr", "multi", "bfs_multi", "many"])
    ap.add_argument("--graphs", type=int, default=512,
                    help="many: Kronecker graphs of scale U{10..16}, edge factor U{8..32}")
    ap.add_argument("--sources", type=int, default=4)
    ap.add_argument("--segments", type=int, default=0,
                    help="also build a host/GPU schedule over this many segments")
    ap.add_argument("--cache", default="autotune_cache.json")
    ap.add_argument("--tree-depth", type=int, default=-1,
                    help="also run the tree microbenchmark (2^0..2^D segments)")
    ap.add_argument("--yaml-dir", default="",
                    help="write the segment/tree profiles here as the reference's YAML")
    args = ap.parse_args(argv)
    hc.check(hc.lib().hc_init(0))
    if args.algorithm == "many":
        graphs = many_graphs(args.graphs)
        sources = [int(x.pick_sources(1)[0]) for x in graphs]
        with hc.GraphBatch(graphs) as b:
            res = tune_knobs(b, "many", sources, log=print)
        print(json.dumps({k: v for k, v in res.items() if k != "trials"}))
        save(args.cache, cache_key(graphs, "many"), res)
        return
    g = hc.HostGraph.generate(args.kind, args.scale, args.degree,
                              weighted=args.algorithm in ("sssp", "multi"))
    n_src = 64 if args.algorithm in ("multi", "bfs_multi") else args.sources
    sources = g.pick_sources(n_src)
    with hc.DeviceGraph(g) as dg:
        res = tune_knobs(dg, args.algorithm, sources, log=print, per_graph=True)
    print(json.dumps({k: v for k, v in res.items() if k != "trials"}))
    save(args.cache, cache_key(g, args.algorithm), res)
    if args.segments and args.algorithm != "multi":
        sched, info = tune_schedule(g, args.algorithm, args.segments)
        print(sched.text())
        print(json.dumps({k: v for k, v in info.items() if k != "profile"}))
        if args.yaml_dir:
            for path in write_yaml_profiles(info["profile"], g, args.yaml_dir):
                print(path)
    if args.tree_depth >= 0 and args.algorithm != "multi":
        tree = profile_tree(g, args.algorithm, args.tree_depth)
        for path in write_yaml_tree(tree, g, args.yaml_dir or "."):
            print(path)


if __name__ == "__main__":
    main()
