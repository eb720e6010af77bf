// Many small graphs, one BFS each, in one persistent launch.
//
// The reference answers one BFS per driver call (bfs_gpu,
// src/kernels/gpu/bfs.cuh:21-95; bfs_do_cpu, src/kernels/cpu/bfs.cuh:48-114).
// For hundreds of graphs of 10^3-10^5 vertices a full-device launch per
// epoch leaves the GPU idle, so here each workgroup runs a whole
// direction-optimizing BFS of one graph, levels separated by workgroup
// barriers only, with the frontier, next-frontier and visited bitmaps of the
// graph in LDS.  Workgroups take graphs from a device queue in decreasing
// edge count (longest job first) until the queue is empty.
//
// Per graph the level loop follows bfs_do_cpu's switching rule exactly
// (push -> pull when scout > edges_to_check / alpha; pull continues while the
// frontier grows or exceeds n / beta), so depths and the number of push and
// pull levels equal the single-graph run; parents are valid BFS parents.
#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/hc_batch.h"
#include "device_common.h"

struct hc_graph_batch_s {
    int32_t num_graphs = 0;
    int64_t total_nodes = 0;
    int64_t total_edges = 0;
    int32_t max_nodes = 0;
    void *dmem = nullptr;  // descriptors | order | offsets | adjacency
    size_t bytes = 0;
    std::vector<int64_t> node_base;   // num_graphs + 1
    std::vector<int64_t> edge_count;  // per graph
    std::vector<int32_t> order;       // graphs by decreasing edge count
    // Device views.
    struct Desc *desc = nullptr;
    int32_t *d_order = nullptr;
    uint32_t *offsets = nullptr;
    int32_t *adj = nullptr;
    // Global bitmaps for graphs above the LDS limit, one slot per workgroup.
    uint32_t *scratch = nullptr;
    int32_t scratch_slots = 0;
    size_t scratch_words = 0;
    int32_t *queue = nullptr;  // work counter + statistics
    int32_t *d_src = nullptr;  // per-call sources
    std::mutex mu;             // one call at a time per batch
};

struct Desc {
    int64_t node_base;  // first vertex of the graph in the batch
    int64_t off_base;   // first offset (node_base + graph index)
    int64_t edge_base;  // first adjacency entry
    int32_t n;
    int32_t pad;
    int64_t m;
};

namespace hc {

// Graphs with at most many.lds_nodes vertices (at most kManyLdsNodesMax:
// 3 x 4096 words = 48 KiB per workgroup) keep their bitmaps in LDS; the LDS
// of a launch is sized to the largest such graph of the batch.

// Launch shape and per-level thresholds of one many-graph launch (knobs
// many.block, many.per_cu, many.light_deg, many.heavy_cap, many.lds_nodes).
struct ManyShape {
    int block;      // threads per workgroup: 256, 512 or 1024
    int light_deg;  // top-down: a thread expands frontier vertices up to this degree
    int heavy_cap;  // heavier ones (at most this many per level) go to whole waves
    int lds_words;  // bitmap words per LDS bitmap (graphs of <= 32x this many vertices)
};

struct ManyStats {
    unsigned long long pushes, pulls, scanned;
};

// Bitmap word read: LDS bitmaps are read directly; global ones bypass the
// per-CU cache, since other waves of the workgroup update them with
// device-memory atomics.
template <bool Global>
__device__ __forceinline__ uint32_t bm_word(const uint32_t *b, int32_t w) {
    if (Global) return __hip_atomic_load(&b[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return b[w];
}

template <bool Global>
__device__ __forceinline__ bool bm_get(const uint32_t *b, int32_t v) {
    return (bm_word<Global>(b, v >> 5) >> (v & 31)) & 1u;
}

// Claims v in the visited bitmap; true for exactly one caller.
template <bool Global>
__device__ __forceinline__ bool bm_claim(uint32_t *b, int32_t v) {
    const uint32_t bit = 1u << (v & 31);
    if (bm_word<Global>(b, v >> 5) & bit) return false;
    return !(atomicOr(&b[v >> 5], bit) & bit);
}

// Block-wide sum of a per-thread 64-bit value (result valid in all threads).
template <int BLOCK>
__device__ __forceinline__ long long many_block_sum(long long v, long long *s_red) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, kWave);
    __syncthreads();
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = v;
    __syncthreads();
    long long t = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / kWave; i++) t += s_red[i];
    return t;
}

// One direction-optimizing BFS of graph d from src by the whole workgroup;
// bm holds three bitmaps of `words` words (LDS, or a global slot when
// Global).  Frontier vertices of degree above light_deg are expanded by
// whole wavefronts, up to heavy_cap of them per level.
template <int BLOCK, bool Global>
__device__ __forceinline__ void bfs_one_graph(const Desc &d, int32_t src,
                              const uint32_t *__restrict__ off,
                              const int32_t *__restrict__ adj, int32_t *__restrict__ depth,
                              int32_t *__restrict__ parents, uint32_t *bm, int words,
                              int alpha, int beta, int light_deg, int heavy_cap,
                              int32_t *heavy, int *s_nheavy, long long *s_red, ManyStats &st) {
    const int tid = threadIdx.x;
    const int n = d.n;
    uint32_t *vis = bm;
    uint32_t *front = bm + words;
    uint32_t *next = bm + 2 * words;
    for (int i = tid; i < 3 * words; i += BLOCK) bm[i] = 0;
    for (int u = tid; u < n; u += BLOCK) {
        depth[u] = -1;
        if (parents) parents[u] = -1;
    }
    __syncthreads();
    if (tid == 0) {
        vis[src >> 5] |= 1u << (src & 31);
        front[src >> 5] |= 1u << (src & 31);
        depth[src] = 0;
        if (parents) parents[src] = src;
    }
    __syncthreads();
    long long edges_to_check = d.m;
    long long scout = (long long)(off[src + 1] - off[src]);
    long long num_nodes = 1, prev_nodes = 0;
    bool pull = false;
    // Initial decision (bfs_do_cpu: num_edges = deg(src)).
    if (scout > edges_to_check / alpha) {
        pull = true;
    } else {
        edges_to_check -= scout;
    }
    long long fsize = 1;
    int level = 0;
    while (fsize > 0) {
        level++;
        long long cnt = 0, sc = 0, scanned = 0;
        if (!pull) {
            // Top-down: threads take frontier words; light vertices are
            // expanded by their thread, heavy ones queued for whole waves.
            if (tid == 0) *s_nheavy = 0;
            __syncthreads();
            for (int w = tid; w < words; w += BLOCK) {
                uint32_t bits = bm_word<Global>(front, w);
                while (bits) {
                    const int32_t u = w * 32 + __ffs(bits) - 1;
                    bits &= bits - 1;
                    const uint32_t b = off[u], e = off[u + 1];
                    if (e - b > (uint32_t)light_deg) {
                        const int slot = atomicAdd(s_nheavy, 1);
                        if (slot < heavy_cap) {
                            heavy[slot] = u;
                            continue;
                        }
                    }
                    scanned += e - b;
                    for (uint32_t i = b; i < e; i++) {
                        const int32_t v = adj[i];
                        if (bm_claim<Global>(vis, v)) {
                            atomicOr(&next[v >> 5], 1u << (v & 31));
                            depth[v] = level;
                            if (parents) parents[v] = u;
                            cnt++;
                            sc += off[v + 1] - off[v];
                        }
                    }
                }
            }
            __syncthreads();
            const int nh = min(*s_nheavy, heavy_cap);
            const int wave = tid / kWave, nwaves = BLOCK / kWave, lane = lane_id();
            for (int h = wave; h < nh; h += nwaves) {
                const int32_t u = heavy[h];
                const uint32_t b = off[u], e = off[u + 1];
                if (lane == 0) scanned += e - b;
                for (uint32_t i = b + lane; i < e; i += kWave) {
                    const int32_t v = adj[i];
                    if (bm_claim<Global>(vis, v)) {
                        atomicOr(&next[v >> 5], 1u << (v & 31));
                        depth[v] = level;
                        if (parents) parents[v] = u;
                        cnt++;
                        sc += off[v + 1] - off[v];
                    }
                }
            }
            st.pushes += (tid == 0);
        } else {
            // Bottom-up: every unvisited vertex takes its first neighbour
            // (CSR order) in the frontier.
            for (int u = tid; u < n; u += BLOCK) {
                if (bm_get<Global>(vis, u)) continue;
                const uint32_t b = off[u], e = off[u + 1];
                for (uint32_t i = b; i < e; i++) {
                    const int32_t v = adj[i];
                    scanned++;
                    if (bm_get<Global>(front, v)) {
                        atomicOr(&vis[u >> 5], 1u << (u & 31));
                        atomicOr(&next[u >> 5], 1u << (u & 31));
                        depth[u] = level;
                        if (parents) parents[u] = v;
                        cnt++;
                        break;
                    }
                }
            }
            st.pulls += (tid == 0);
        }
        const long long q = many_block_sum<BLOCK>(cnt, s_red);
        const long long scq = many_block_sum<BLOCK>(sc, s_red);
        st.scanned += (unsigned long long)scanned;
        // Next level's frontier; clear the old one for reuse.
        uint32_t *t = front;
        front = next;
        next = t;
        for (int i = tid; i < words; i += BLOCK) next[i] = 0;
        fsize = q;
        // bfs_do_cpu's switching rule, evaluated identically by all threads.
        if (!pull) {
            scout = scq;
            if (q > 0) {
                if (scout > edges_to_check / alpha) {
                    pull = true;
                    num_nodes = q;
                } else {
                    edges_to_check -= scout;
                }
            }
        } else {
            prev_nodes = num_nodes;
            num_nodes = q;
            if (!(q > prev_nodes || q > n / beta)) {
                scout = 1;
                if (q > 0) {
                    if (scout > edges_to_check / alpha) {
                        num_nodes = q;
                    } else {
                        pull = false;
                        edges_to_check -= scout;
                    }
                }
            }
        }
        __syncthreads();
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void bfs_many_kernel(
    const Desc *__restrict__ desc, const int32_t *__restrict__ order,
    const int32_t *__restrict__ sources, int num_graphs,
    const uint32_t *__restrict__ offsets, const int32_t *__restrict__ adjacency,
    int32_t *__restrict__ depth, int32_t *__restrict__ parents, uint32_t *slots,
    int slot_words, int *queue, unsigned long long *stats, int alpha, int beta,
    ManyShape shape) {
    extern __shared__ uint32_t lds_bm[];
    __shared__ int32_t heavy[kManyHeavyCapMax];
    __shared__ int s_nheavy, s_g;
    __shared__ long long s_red[BLOCK / kWave];
    ManyStats st{0, 0, 0};
    const int lds_nodes = shape.lds_words * 32;
    for (;;) {
        if (threadIdx.x == 0) s_g = atomicAdd(queue, 1);
        __syncthreads();
        const int k = s_g;
        __syncthreads();
        if (k >= num_graphs) break;
        const int gi = order[k];
        const Desc d = desc[gi];
        const int words = (d.n + 31) / 32;
        int32_t *par = parents ? parents + d.node_base : nullptr;
        if (d.n <= lds_nodes)
            bfs_one_graph<BLOCK, false>(d, sources[gi], offsets + d.off_base,
                                        adjacency + d.edge_base, depth + d.node_base, par, lds_bm,
                                        words, alpha, beta, shape.light_deg, shape.heavy_cap,
                                        heavy, &s_nheavy, s_red, st);
        else
            bfs_one_graph<BLOCK, true>(d, sources[gi], offsets + d.off_base,
                                       adjacency + d.edge_base, depth + d.node_base, par,
                                       slots + (size_t)blockIdx.x * 3 * slot_words, words, alpha,
                                       beta, shape.light_deg, shape.heavy_cap, heavy, &s_nheavy,
                                       s_red, st);
    }
    // Levels are counted once per graph (thread 0); edges by every thread.
    const long long sc = many_block_sum<BLOCK>((long long)st.scanned, s_red);
    if (threadIdx.x == 0) {
        atomicAdd(&stats[0], st.pushes);
        atomicAdd(&stats[1], st.pulls);
        atomicAdd(&stats[2], (unsigned long long)sc);
    }
}

// Clears the work counter and the statistics words of a call (a kernel, not
// hipMemsetAsync, so that the launch sequence can be captured in a graph).
__global__ void many_clear_kernel(int32_t *queue) {
    if (threadIdx.x < 16) queue[threadIdx.x] = 0;
}

}  // namespace hc

using namespace hc;

namespace {

struct ManyLaunch {
    int32_t *d_src;        // device sources (num_graphs)
    uint32_t *slots;       // bitmap slots, slot_words x 3 per workgroup
    int slot_words;
    int grid;
    int32_t *queue;        // 16 words: counter, statistics
    ManyShape shape;
};

// Launch shape of a batch under the current knobs: LDS bitmaps sized to the
// largest graph that fits the many.lds_nodes limit, global bitmap slots of
// slot_words words for the graphs above it (0 when none is), and the
// workgroups (at most one per graph) that the device holds at once with
// that LDS, or many.per_cu per CU when set.
ManyShape many_shape(const hc_graph_batch_s *b, int *grid, int *slot_words) {
    const Tuning &t = tuning();
    ManyShape sh{t.many_block, t.many_light_deg, t.many_heavy_cap, 0};
    const int limit = std::min(t.many_lds_nodes, kManyLdsNodesMax);
    int fit = 0;
    for (int i = 0; i < b->num_graphs; i++) {
        const int n = (int)(b->node_base[i + 1] - b->node_base[i]);
        if (n <= limit) fit = std::max(fit, n);
    }
    sh.lds_words = (fit + 31) / 32;
    *slot_words = b->max_nodes > sh.lds_words * 32 ? (b->max_nodes + 31) / 32 : 0;
    const size_t lds = sizeof(uint32_t) * 3 * (size_t)sh.lds_words;
    int per_cu = t.many_per_cu;
    if (per_cu <= 0) {
        const void *k = sh.block == 256   ? reinterpret_cast<const void *>(&bfs_many_kernel<256>)
                        : sh.block == 512 ? reinterpret_cast<const void *>(&bfs_many_kernel<512>)
                                          : reinterpret_cast<const void *>(&bfs_many_kernel<1024>);
        per_cu = resident_grid(k, sh.block, lds, INT32_MAX) / std::max(1, device_state().cus);
    }
    *grid = std::max(1, std::min(b->num_graphs, device_state().cus * std::max(1, per_cu)));
    return sh;
}

// Enqueues one persistent launch; with stats, waits for it and fills them
// (ms: device time between events when `ev` are given, else the host-observed
// time of the launch).
int launch_many(hc_graph_batch_s *b, const ManyLaunch &L, const int32_t *sources,
                int32_t *depth, int32_t *parents, int alpha, int beta, hc_run_stats_t *stats,
                hipStream_t s, bool events) {
    HC_HIP(hipMemcpyAsync(L.d_src, sources, sizeof(int32_t) * b->num_graphs,
                          hipMemcpyHostToDevice, s));
    many_clear_kernel<<<1, 64, 0, s>>>(L.queue);
    HC_HIP(hipGetLastError());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (events) {
        HC_HIP(hipEventCreate(&e0));
        HC_HIP(hipEventCreate(&e1));
        HC_HIP(hipEventRecord(e0, s));
    }
    const auto t0 = std::chrono::steady_clock::now();
    const size_t lds = sizeof(uint32_t) * 3 * (size_t)L.shape.lds_words;
    unsigned long long *st_words = reinterpret_cast<unsigned long long *>(L.queue + 2);
#define HC_MANY_LAUNCH(B)                                                                      \
    bfs_many_kernel<B><<<L.grid, B, lds, s>>>(b->desc, b->d_order, L.d_src, b->num_graphs,     \
                                              b->offsets, b->adj, depth, parents, L.slots,     \
                                              L.slot_words, L.queue, st_words, alpha, beta,    \
                                              L.shape)
    if (L.shape.block == 256)
        HC_MANY_LAUNCH(256);
    else if (L.shape.block == 512)
        HC_MANY_LAUNCH(512);
    else
        HC_MANY_LAUNCH(1024);
#undef HC_MANY_LAUNCH
    HC_HIP(hipGetLastError());
    if (!stats && !events) return HC_OK;
    float ms = 0;
    if (events) {
        HC_HIP(hipEventRecord(e1, s));
        HC_HIP(hipEventSynchronize(e1));
        HC_HIP(hipEventElapsedTime(&ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
    }
    unsigned long long st[3] = {0, 0, 0};
    HC_HIP(hipMemcpyAsync(st, L.queue + 2, sizeof(st), hipMemcpyDeviceToHost, s));
    HC_HIP(hipStreamSynchronize(s));
    if (!events)
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                               t0).count();
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = ms;
        stats->push_epochs = (int32_t)st[0];
        stats->pull_epochs = (int32_t)st[1];
        stats->epochs = (int32_t)(st[0] + st[1]);
        stats->launches = 1;
        stats->reached = -1;
        // Offsets (2 x 4 B per visited-check), ids of scanned edges, depth
        // and parent initialisation and writes.
        stats->alg_bytes = 4.0 * (double)st[2] + 16.0 * (double)b->total_nodes;
        stats->kernel_ms = ms;
        stats->kernel_launches = 1;
    }
    return HC_OK;
}

int check_many_args(hc_graph_batch_s *b, const int32_t *sources, const int32_t *depth) {
    HC_CHECK_ARG(b && sources && depth, "NULL argument");
    for (int i = 0; i < b->num_graphs; i++) {
        const int32_t n = (int32_t)(b->node_base[i + 1] - b->node_base[i]);
        HC_CHECK_ARG(sources[i] >= 0 && sources[i] < n, "graph %d: source %d outside [0,%d)", i,
                     sources[i], n);
    }
    return HC_OK;
}

void free_batch(hc_graph_batch_s *b) {
    if (!b) return;
    if (b->dmem) hipFree(b->dmem);
    if (b->scratch) hipFree(b->scratch);
    if (b->queue) hipFree(b->queue);
    if (b->d_src) hipFree(b->d_src);
    delete b;
}

// Scratch of hc_bfs_many_noalloc, for callers that cannot allocate device
// memory: the work counter and statistics words (padded to 256 bytes), the
// sources of a call, and the global bitmap slots (three bitmaps per
// workgroup) of graphs above the LDS limit.  Each GPU holds its own copy
// from the first use of this file's kernels until the process ends; calls
// that overlap on one GPU share it and race.
constexpr int kNoAllocMaxGraphs = 1 << 16;
constexpr size_t kNoAllocHeadBytes = 256;
constexpr size_t kNoAllocSrcBytes = sizeof(int32_t) * kNoAllocMaxGraphs;
constexpr size_t kNoAllocSlotBytes = size_t(32) << 20;
constexpr size_t kNoAllocBytes = kNoAllocHeadBytes + kNoAllocSrcBytes + kNoAllocSlotBytes;
constexpr int kNoAllocMaxDevices = 64;

}  // namespace

static __device__ __align__(256) uint32_t scratch[kNoAllocBytes / sizeof(uint32_t)];

namespace {

// Device address of `scratch` on the current GPU, looked up once per device.
// The host-side name `scratch` is not a device address and is never passed
// to a kernel or a copy.
int noalloc_scratch(char **out) {
    static std::mutex mu;
    static char *ptrs[kNoAllocMaxDevices] = {};
    int dev = 0;
    HC_HIP(hipGetDevice(&dev));
    HC_CHECK_ARG(dev >= 0 && dev < kNoAllocMaxDevices, "device %d above the %d supported", dev,
                 kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[dev]) {
        void *p = nullptr;
        HC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HC_CHECK_ARG(p != nullptr, "no device address for the scratch array");
        ptrs[dev] = static_cast<char *>(p);
    }
    *out = ptrs[dev];
    return HC_OK;
}

}  // namespace

extern "C" {

int hc_graph_batch_upload(const hc_host_graph_t *graphs, int32_t num_graphs,
                          hc_graph_batch_t *out) {
    HC_CHECK_ARG(graphs && out, "NULL argument");
    HC_CHECK_ARG(num_graphs >= 1, "num_graphs must be >= 1");
    HC_TRY(ensure_init());
    *out = nullptr;
    auto *b = new (std::nothrow) hc_graph_batch_s();
    if (!b) return fail(HC_ERR_OOM, "host allocation failed");
    b->num_graphs = num_graphs;
    b->node_base.assign(num_graphs + 1, 0);
    b->edge_count.resize(num_graphs);
    std::vector<Desc> desc(num_graphs);
    int64_t nodes = 0, edges = 0, offs = 0;
    for (int i = 0; i < num_graphs; i++) {
        const hc_host_graph_t &g = graphs[i];
        if (!g.index || g.num_nodes <= 0 || (g.num_edges > 0 && !g.neighbors) ||
            g.index[0] != 0 || g.index[g.num_nodes] != g.num_edges ||
            g.num_edges >= (int64_t(1) << 32)) {
            free_batch(b);
            return fail(HC_ERR_INVALID, "graph %d: empty, inconsistent or above 2^32 edges", i);
        }
        for (int32_t u = 0; u < g.num_nodes; u++)
            if (g.index[u + 1] < g.index[u]) {
                free_batch(b);
                return fail(HC_ERR_INVALID, "graph %d: index not monotone at row %d", i, u);
            }
        for (int64_t e = 0; e < g.num_edges; e++)
            if (g.neighbors[e] < 0 || g.neighbors[e] >= g.num_nodes) {
                free_batch(b);
                return fail(HC_ERR_INVALID, "graph %d: neighbour %lld = %d outside [0,%d)", i,
                            (long long)e, g.neighbors[e], g.num_nodes);
            }
        desc[i] = Desc{nodes, offs, edges, g.num_nodes, 0, g.num_edges};
        b->node_base[i] = nodes;
        b->edge_count[i] = g.num_edges;
        b->max_nodes = std::max(b->max_nodes, g.num_nodes);
        nodes += g.num_nodes;
        offs += g.num_nodes + 1;
        edges += g.num_edges;
    }
    b->node_base[num_graphs] = nodes;
    b->total_nodes = nodes;
    b->total_edges = edges;
    b->order.resize(num_graphs);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        return b->edge_count[x] > b->edge_count[y];
    });
    // One allocation: descriptors | order | offsets (u32) | adjacency (i32).
    size_t off_order = sizeof(Desc) * num_graphs;
    size_t off_offs = (off_order + sizeof(int32_t) * num_graphs + 255) & ~size_t(255);
    size_t off_adj = (off_offs + sizeof(uint32_t) * offs + 255) & ~size_t(255);
    b->bytes = off_adj + sizeof(int32_t) * std::max<int64_t>(edges, 1);
    if (hipMalloc(&b->dmem, b->bytes) != hipSuccess) {
        b->dmem = nullptr;
        free_batch(b);
        return fail(HC_ERR_OOM, "device allocation of %zu bytes failed", b->bytes);
    }
    char *base = static_cast<char *>(b->dmem);
    b->desc = reinterpret_cast<Desc *>(base);
    b->d_order = reinterpret_cast<int32_t *>(base + off_order);
    b->offsets = reinterpret_cast<uint32_t *>(base + off_offs);
    b->adj = reinterpret_cast<int32_t *>(base + off_adj);
    std::vector<uint32_t> o32;
    int rc = HC_OK;
    if (hipMemcpy(b->desc, desc.data(), sizeof(Desc) * num_graphs, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(b->d_order, b->order.data(), sizeof(int32_t) * num_graphs,
                  hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(HC_ERR_HIP, "descriptor upload failed");
    for (int i = 0; rc == HC_OK && i < num_graphs; i++) {
        const hc_host_graph_t &g = graphs[i];
        o32.assign(g.index, g.index + g.num_nodes + 1);
        if (hipMemcpy(b->offsets + desc[i].off_base, o32.data(),
                      sizeof(uint32_t) * o32.size(), hipMemcpyHostToDevice) != hipSuccess ||
            (g.num_edges > 0 &&
             hipMemcpy(b->adj + desc[i].edge_base, g.neighbors, sizeof(int32_t) * g.num_edges,
                       hipMemcpyHostToDevice) != hipSuccess))
            rc = fail(HC_ERR_HIP, "graph %d upload failed", i);
    }
    if (rc == HC_OK && (hipMalloc(&b->queue, 64) != hipSuccess ||
                        hipMalloc(&b->d_src, sizeof(int32_t) * num_graphs) != hipSuccess))
        rc = fail(HC_ERR_OOM, "queue allocation failed");
    if (rc != HC_OK) {
        free_batch(b);
        return rc;
    }
    *out = b;
    return HC_OK;
}

int hc_graph_batch_free(hc_graph_batch_t b) {
    free_batch(b);
    return HC_OK;
}

int hc_graph_batch_info(hc_graph_batch_t b, int32_t *num_graphs, int64_t *total_nodes,
                        int64_t *total_edges, size_t *device_bytes) {
    HC_CHECK_ARG(b, "NULL batch");
    if (num_graphs) *num_graphs = b->num_graphs;
    if (total_nodes) *total_nodes = b->total_nodes;
    if (total_edges) *total_edges = b->total_edges;
    if (device_bytes) *device_bytes = b->bytes;
    return HC_OK;
}

int hc_graph_batch_node_offsets(hc_graph_batch_t b, int64_t *out) {
    HC_CHECK_ARG(b && out, "NULL argument");
    std::memcpy(out, b->node_base.data(), sizeof(int64_t) * (b->num_graphs + 1));
    return HC_OK;
}

int hc_bfs_many(hc_graph_batch_t b, const int32_t *sources, int32_t *depth,
                int32_t *parents, int32_t alpha, int32_t beta, hc_run_stats_t *stats,
                hc_stream_t stream) {
    HC_TRY(check_many_args(b, sources, depth));
    if (alpha <= 0) alpha = 15;
    if (beta <= 0) beta = 18;
    HC_TRY(ensure_init());
    std::lock_guard<std::mutex> lk(b->mu);
    hipStream_t s = pick_stream(stream);
    int grid = 1, slot_words = 0;
    const ManyShape shape = many_shape(b, &grid, &slot_words);
    if (slot_words > 0) {
        const size_t need = 3 * (size_t)slot_words * grid;
        if (b->scratch_slots < grid || b->scratch_words < need) {
            if (b->scratch) hipFree(b->scratch);
            b->scratch = nullptr;
            b->scratch_slots = 0;
            b->scratch_words = 0;
            HC_HIP(hipMalloc(&b->scratch, sizeof(uint32_t) * need));
            b->scratch_slots = grid;
            b->scratch_words = need;
        }
    }
    const ManyLaunch L{b->d_src, b->scratch, slot_words, grid, b->queue, shape};
    return launch_many(b, L, sources, depth, parents, alpha, beta, stats, s, true);
}

int hc_bfs_many_noalloc(hc_graph_batch_t b, const int32_t *sources, int32_t *depth,
                        int32_t *parents, int32_t alpha, int32_t beta, hc_run_stats_t *stats,
                        hc_stream_t stream) {
    HC_TRY(check_many_args(b, sources, depth));
    HC_CHECK_ARG(b->num_graphs <= kNoAllocMaxGraphs,
                 "%d graphs exceed the %d that the allocation-free scratch holds",
                 b->num_graphs, kNoAllocMaxGraphs);
    if (alpha <= 0) alpha = 15;
    if (beta <= 0) beta = 18;
    HC_TRY(ensure_init());
    char *base = nullptr;
    HC_TRY(noalloc_scratch(&base));
    std::lock_guard<std::mutex> lk(b->mu);
    hipStream_t s = pick_stream(stream);
    int grid = 1, slot_words = 0;
    const ManyShape shape = many_shape(b, &grid, &slot_words);
    if (slot_words > 0) {
        // Fewer workgroups when the slots of a full grid do not fit: the
        // persistent kernel takes graphs from its queue until none is left.
        const size_t per_slot = sizeof(uint32_t) * 3 * (size_t)slot_words;
        const size_t fit = kNoAllocSlotBytes / per_slot;
        HC_CHECK_ARG(fit >= 1,
                     "a graph of %d vertices needs %zu bytes of bitmap scratch, above the "
                     "%zu-byte maximum of the allocation-free entry point",
                     b->max_nodes, per_slot, kNoAllocSlotBytes);
        grid = (int)std::min<size_t>((size_t)grid, fit);
    }
    const ManyLaunch L{reinterpret_cast<int32_t *>(base + kNoAllocHeadBytes),
                       reinterpret_cast<uint32_t *>(base + kNoAllocHeadBytes + kNoAllocSrcBytes),
                       slot_words, grid, reinterpret_cast<int32_t *>(base), shape};
    return launch_many(b, L, sources, depth, parents, alpha, beta, stats, s, false);
}

}  // extern "C"
