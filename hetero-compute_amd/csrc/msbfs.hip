// Breadth-first depths from up to 64 sources at once, bit-parallel.
//
// The reference answers one BFS per driver call (bfs_gpu,
// src/kernels/gpu/bfs.cuh:21-95; bfs_do_cpu, src/kernels/cpu/bfs.cuh:48-114).
// Here the 64 searches share every sweep: each vertex keeps a 64-bit mask of
// the sources that have reached it (seen) and, per level, the mask of the
// sources whose frontier it is on.  Level L's frontier bit q of v is set
// exactly when v's depth from source q is L, so the depths are the unique
// BFS levels, identical to single-source runs.  A level costs one pass over
// the edges that matter with an 8-byte gather per edge, instead of one pass
// per source.
//
// Levels alternate between two directions, chosen from the frontier's edge
// count as bfs_do_cpu chooses between push and pull:
//  * push: the frontier list's vertices OR their frontier mask, less the
//    target's seen mask, into the target's next mask (an atomic only when
//    it adds a bit), and the first to set a target's mask appends it to the
//    next list;
//  * pull: every vertex not yet seen from all sources ORs its neighbours'
//    frontier masks (lanes stop once they hold every missing bit), over the
//    degree-bucketed plan, and records the bits it did not have.
// A commit pass over the next list then sets the depth entries of the new
// bits and folds them into seen.  Frontier masks and lists rotate over three
// buffers: level L reads buffer L-1, writes buffer L and clears buffer L-2
// by its list, so no pass clears a whole array.
//
// Depths are the int32 matrix of the batched SSSP interface, vertex-major
// (depth[v * 64 + q]), filled with HC_INF_DIST first.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/hc_batch.h"
#include "device_common.h"

namespace hc {

constexpr int kMsQ = 64;
constexpr int kMsLight = 32;      // push: a thread expands frontier vertices up to this degree
constexpr int kMsHeavyCap = 512;  // heavier ones (per workgroup pass) go to whole wavefronts

// Per-level counters of one rotation slot: list length, its edge count
// (sum of degrees), edges examined, depth entries written.
struct MsCount {
    uint32_t len;
    uint32_t pad;
    unsigned long long deg, edges, bits;
};

struct MsBufs {
    uint64_t *seen;
    uint64_t *fr[3];
    int32_t *q[3];
    MsCount *cnt;  // [3]
    int32_t *depth;
    int32_t n;
    uint64_t full;  // bits of the sources in use
};

struct MsSeeds {
    int32_t v[kMsQ];
    uint64_t mask[kMsQ];
    int32_t k;
};

__device__ __forceinline__ uint64_t or_xor_shfl(uint64_t v, int s, int width) {
    return v | (uint64_t)__shfl_xor((unsigned long long)v, s, width);
}

template <int G>
__device__ __forceinline__ uint64_t group_or(uint64_t v) {
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) v = or_xor_shfl(v, s, G);
    return v;
}

// OR over the 256-thread workgroup; valid in thread 0.
__device__ __forceinline__ uint64_t block_or(uint64_t v, uint64_t *s_part) {
    v = group_or<kWave>(v);
    __syncthreads();
    if (lane_id() == 0) s_part[threadIdx.x / kWave] = v;
    __syncthreads();
    uint64_t t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWavesPerBlock; i++) t |= s_part[i];
    return t;
}

// Depth matrix to HC_INF_DIST, masks and seen to zero.
__global__ __launch_bounds__(kBlock) void ms_fill_kernel(MsBufs b) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    const size_t cells = (size_t)b.n * kMsQ;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < cells; i += stride)
        b.depth[i] = HC_INF_DIST;
    for (size_t v = blockIdx.x * (size_t)kBlock + threadIdx.x; v < (size_t)b.n; v += stride) {
        b.seen[v] = 0;
        b.fr[0][v] = 0;
        b.fr[1][v] = 0;
        b.fr[2][v] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 3) b.cnt[threadIdx.x] = MsCount{0, 0, 0, 0, 0};
}

// Level 0: the distinct source vertices with their masks.
template <typename OffT>
__global__ void ms_seed_kernel(MsBufs b, GraphView<OffT> g, MsSeeds s) {
    const int t = threadIdx.x;
    if (t < s.k) {
        const int32_t v = s.v[t];
        const uint64_t m = s.mask[t];
        b.seen[v] = m;
        b.fr[0][v] = m;
        b.q[0][t] = v;
        for (int q = 0; q < kMsQ; q++)
            if ((m >> q) & 1) b.depth[(size_t)v * kMsQ + q] = 0;
        atomicAdd(&b.cnt[0].deg, (unsigned long long)(g.rend(v) - g.rbeg(v)));
    }
    if (t == 0) b.cnt[0].len = (uint32_t)s.k;
}

// Clears the masks of level L-2 (buffer (L+1) % 3) by its list.
__device__ __forceinline__ void ms_clear_old(const MsBufs &b, int L) {
    const int o = (L + 1) % 3;
    const unsigned long long len = b.cnt[o].len;
    uint64_t *f = b.fr[o];
    const int32_t *q = b.q[o];
    for (unsigned long long i = blockIdx.x * (unsigned long long)kBlock + threadIdx.x; i < len;
         i += (unsigned long long)gridDim.x * kBlock)
        f[q[i]] = 0;
}

// Edges [i0, e) with stride `step` of a frontier vertex with mask fu: four
// targets per trip, their ids, then their seen and next masks, loaded
// before any is used.
template <typename OffT>
__device__ __forceinline__ void ms_push_edges(const MsBufs &b, const GraphView<OffT> &g,
                                              uint64_t *nxt, int32_t *nq, MsCount &c,
                                              uint64_t fu, OffT i0, OffT e, OffT step,
                                              unsigned long long &deg_acc,
                                              unsigned long long &edges) {
    for (OffT i = i0; i < e; i += 4 * step) {
        int32_t v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const OffT k = i + OffT(r) * step;
            v[r] = k < e ? ldnt(&g.adj[k]) : -1;
        }
        uint64_t bits[4], have[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            bits[r] = v[r] >= 0 ? fu & ~b.seen[v[r]] : 0;
            have[r] = v[r] >= 0 ? nxt[v[r]] : 0;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            edges += v[r] >= 0;
            bool first = false;
            // An atomic only when it adds a bit the next mask lacks.
            if (bits[r] & ~have[r])
                first = atomicOr((unsigned long long *)&nxt[v[r]],
                                 (unsigned long long)bits[r]) == 0;
            if (first) deg_acc += (unsigned long long)(g.rend(v[r]) - g.rbeg(v[r]));
            wave_append(nq, &c.len, first, v[r]);
        }
    }
}

// Push level L over the list of level L-1.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void ms_push_kernel(MsBufs b, GraphView<OffT> g, int L) {
    __shared__ int32_t heavy[kMsHeavyCap];
    __shared__ int s_nheavy;
    const int cu = (L + 2) % 3, nx = L % 3;
    const uint64_t *cur = b.fr[cu];
    uint64_t *nxt = b.fr[nx];
    const int32_t *list = b.q[cu];
    int32_t *nq = b.q[nx];
    MsCount &c = b.cnt[nx];
    ms_clear_old(b, L);
    // Counters of buffer nx were reset by the previous commit; the list
    // length read here is level L-1's, final since its commit.
    const unsigned long long len = b.cnt[cu].len;
    unsigned long long deg_acc = 0, edges = 0;
    const unsigned long long stride = (unsigned long long)gridDim.x * kBlock;
    // Rounds of kBlock list entries per workgroup, so the heavy list is
    // per round and every thread reaches each barrier.
    for (unsigned long long base = blockIdx.x * (unsigned long long)kBlock; base < len;
         base += stride) {
        if (threadIdx.x == 0) s_nheavy = 0;
        __syncthreads();
        const unsigned long long i = base + threadIdx.x;
        if (i < len) {
            const int32_t u = list[i];
            const OffT ub = g.rbeg(u), ue = g.rend(u);
            bool self = true;
            if (ue - ub > OffT(kMsLight)) {
                const int slot = atomicAdd(&s_nheavy, 1);
                if (slot < kMsHeavyCap) {
                    heavy[slot] = u;
                    self = false;
                }
            }
            if (self)
                ms_push_edges<OffT>(b, g, nxt, nq, c, cur[u], ub, ue, OffT(1), deg_acc, edges);
        }
        __syncthreads();
        const int nh = min(s_nheavy, kMsHeavyCap);
        for (int h = threadIdx.x / kWave; h < nh; h += kWavesPerBlock) {
            const int32_t u = heavy[h];
            ms_push_edges<OffT>(b, g, nxt, nq, c, cur[u], g.rbeg(u) + OffT(lane_id()), g.rend(u),
                                OffT(kWave), deg_acc, edges);
        }
        __syncthreads();
    }
    const unsigned long long d = block_sum_u64(deg_acc);
    const unsigned long long ed = block_sum_u64(edges);
    if (threadIdx.x == 0) {
        if (d) atomicAdd(&c.deg, d);
        if (ed) atomicAdd(&c.edges, ed);
    }
}

// One row (or hub chunk) of the pull sweep by a group of G lanes: OR of the
// neighbours' frontier masks, four edges per lane per trip (ids, then the
// four gathers, in flight together), each lane stopping once it holds every
// bit the row is missing.
template <int G, typename OffT>
__device__ __forceinline__ uint64_t ms_row_or(const GraphView<OffT> &g, const uint64_t *cur,
                                              OffT b0, OffT e, int lane, uint64_t need,
                                              unsigned long long &edges) {
    uint64_t acc = 0;
    for (OffT i = b0 + OffT(lane); i < e; i += OffT(4 * G)) {
        int32_t v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const OffT k = i + OffT(r * G);
            v[r] = k < e ? ldnt(&g.adj[k]) : -1;
        }
        uint64_t m[4];
#pragma unroll
        for (int r = 0; r < 4; r++) m[r] = v[r] >= 0 ? cur[v[r]] : 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            acc |= m[r];
            edges += v[r] >= 0;
        }
        if ((acc & need) == need) break;
    }
    return acc;
}

// Pull level L over the plan (all rows of the graph).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void ms_pull_kernel(MsBufs b, GraphView<OffT> g, PlanDev p,
                                                         int L) {
    __shared__ uint64_t s_or[kWavesPerBlock];
    const int cu = (L + 2) % 3, nx = L % 3;
    const uint64_t *cur = b.fr[cu];
    uint64_t *nxt = b.fr[nx];
    int32_t *nq = b.q[nx];
    MsCount &c = b.cnt[nx];
    ms_clear_old(b, L);
    unsigned long long deg_acc = 0, edges = 0;
    const int total = p.unit_start[kNumClasses];
    for (int unit = blockIdx.x; unit < total; unit += gridDim.x) {
        int cls = 0;
        while (unit >= p.unit_start[cls + 1]) cls++;
        const int k = unit - p.unit_start[cls];
        if (cls <= 1) {
            // A hub chunk or a whole row per workgroup (unit-uniform).
            int32_t v;
            OffT rb, re;
            if (cls == 0) {
                const HubChunk h = p.chunks[k];
                v = h.u;
                rb = OffT(h.b);
                re = OffT(h.e);
            } else {
                v = p.list[1] ? p.list[1][k] : p.beg[1] + k;
                rb = g.rbeg(v);
                re = g.rend(v);
            }
            const uint64_t sv = b.seen[v];
            const uint64_t need = b.full & ~sv;
            uint64_t acc = 0;
            if (need) acc = ms_row_or<kBlock, OffT>(g, cur, rb, re, threadIdx.x, need, edges);
            const uint64_t got = block_or(acc, s_or) & need;
            bool first = false;
            if (threadIdx.x == 0 && got) {
                if (cls == 0) {
                    first = atomicOr((unsigned long long *)&nxt[v], (unsigned long long)got) == 0;
                } else {
                    nxt[v] = got;
                    first = true;
                }
                if (first) deg_acc += (unsigned long long)(g.rend(v) - g.rbeg(v));
            }
            if (threadIdx.x < kWave) wave_append(nq, &c.len, first, v);
            __syncthreads();
            continue;
        }
        // Rows of classes 2-5: G lanes per row, kBlock / G rows per unit.
        int G = cls == 2 ? 64 : cls == 3 ? 16 : cls == 4 ? 4 : 1;
        const int per = kBlock / G;
        const int idx = k * per + (int)threadIdx.x / G;
        const int li = (int)threadIdx.x % G;
        const bool ok = idx < p.cnt[cls];
        int32_t v = 0;
        uint64_t need = 0, acc = 0;
        if (ok) {
            v = p.list[cls] ? p.list[cls][idx] : p.beg[cls] + idx;
            need = b.full & ~b.seen[v];
            if (need) {
                const OffT rb = g.rbeg(v), re = g.rend(v);
                if (cls == 2)
                    acc = ms_row_or<64, OffT>(g, cur, rb, re, li, need, edges);
                else if (cls == 3)
                    acc = ms_row_or<16, OffT>(g, cur, rb, re, li, need, edges);
                else if (cls == 4)
                    acc = ms_row_or<4, OffT>(g, cur, rb, re, li, need, edges);
                else
                    acc = ms_row_or<1, OffT>(g, cur, rb, re, li, need, edges);
            }
        }
        if (cls == 2)
            acc = group_or<64>(acc);
        else if (cls == 3)
            acc = group_or<16>(acc);
        else if (cls == 4)
            acc = group_or<4>(acc);
        const uint64_t got = acc & need;
        const bool first = ok && li == 0 && got != 0;
        if (first) {
            nxt[v] = got;  // the only group on this row
            deg_acc += (unsigned long long)(g.rend(v) - g.rbeg(v));
        }
        wave_append(nq, &c.len, first, v);
    }
    const unsigned long long d = block_sum_u64(deg_acc);
    const unsigned long long ed = block_sum_u64(edges);
    if (threadIdx.x == 0) {
        if (d) atomicAdd(&c.deg, d);
        if (ed) atomicAdd(&c.edges, ed);
    }
}

// Pull level L over the column blocks of the graph (see colblock.hip):
// units (block k, plan unit) in block-major order, so the grid gathers from
// one window of 2^shift frontier masks at a time.  Plan classes are by mean
// slice degree and have no hub chunks.  A row's bits may come from several
// blocks: they are added with atomicOr (only when the next mask lacks one),
// and the call that made the mask non-zero lists the row.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void ms_pull_cb_kernel(MsBufs b, GraphView<OffT> g,
                                                            GraphView<OffT> cb, int K,
                                                            PlanDev p, int L) {
    __shared__ uint64_t s_or[kWavesPerBlock];
    const int cu = (L + 2) % 3, nx = L % 3;
    const uint64_t *cur = b.fr[cu];
    uint64_t *nxt = b.fr[nx];
    int32_t *nq = b.q[nx];
    MsCount &c = b.cnt[nx];
    ms_clear_old(b, L);
    unsigned long long deg_acc = 0, edges = 0;
    const int U = p.unit_start[kNumClasses];
    const int total = U * K;
    for (int unit = blockIdx.x; unit < total; unit += gridDim.x) {
        const int blk = unit / U;
        const int pu = unit - blk * U;
        GraphView<OffT> gk = cb;
        gk.off = cb.off + (size_t)blk * ((size_t)g.n + 1);
        int cls = 0;
        while (pu >= p.unit_start[cls + 1]) cls++;
        const int k = pu - p.unit_start[cls];
        if (cls <= 1) {
            // A workgroup per row slice (no hub chunks in this plan).
            const int32_t v = p.list[1] ? p.list[1][k] : p.beg[1] + k;
            const uint64_t need = b.full & ~b.seen[v];
            uint64_t acc = 0;
            if (need)
                acc = ms_row_or<kBlock, OffT>(gk, cur, gk.rbeg(v), gk.rend(v), threadIdx.x, need,
                                              edges);
            const uint64_t got = block_or(acc, s_or) & need;
            bool first = false;
            if (threadIdx.x == 0 && (got & ~nxt[v])) {
                first = atomicOr((unsigned long long *)&nxt[v], (unsigned long long)got) == 0;
                if (first) deg_acc += (unsigned long long)(g.rend(v) - g.rbeg(v));
            }
            if (threadIdx.x < kWave) wave_append(nq, &c.len, first, v);
            __syncthreads();
            continue;
        }
        const int G = cls == 2 ? 64 : cls == 3 ? 16 : cls == 4 ? 4 : 1;
        const int per = kBlock / G;
        const int idx = k * per + (int)threadIdx.x / G;
        const int li = (int)threadIdx.x % G;
        const bool ok = idx < p.cnt[cls];
        int32_t v = 0;
        uint64_t need = 0, acc = 0;
        if (ok) {
            v = p.list[cls] ? p.list[cls][idx] : p.beg[cls] + idx;
            need = b.full & ~b.seen[v];
            if (need) {
                const OffT rb = gk.rbeg(v), re = gk.rend(v);
                if (cls == 2)
                    acc = ms_row_or<64, OffT>(gk, cur, rb, re, li, need, edges);
                else if (cls == 3)
                    acc = ms_row_or<16, OffT>(gk, cur, rb, re, li, need, edges);
                else if (cls == 4)
                    acc = ms_row_or<4, OffT>(gk, cur, rb, re, li, need, edges);
                else
                    acc = ms_row_or<1, OffT>(gk, cur, rb, re, li, need, edges);
            }
        }
        if (cls == 2)
            acc = group_or<64>(acc);
        else if (cls == 3)
            acc = group_or<16>(acc);
        else if (cls == 4)
            acc = group_or<4>(acc);
        const uint64_t got = acc & need;
        bool first = false;
        if (ok && li == 0 && (got & ~nxt[v])) {
            first = atomicOr((unsigned long long *)&nxt[v], (unsigned long long)got) == 0;
            if (first) deg_acc += (unsigned long long)(g.rend(v) - g.rbeg(v));
        }
        wave_append(nq, &c.len, first, v);
    }
    const unsigned long long d = block_sum_u64(deg_acc);
    const unsigned long long ed = block_sum_u64(edges);
    if (threadIdx.x == 0) {
        if (d) atomicAdd(&c.deg, d);
        if (ed) atomicAdd(&c.edges, ed);
    }
}

// Commit of level L: depth entries of the new bits, seen updated; one
// wavefront per listed vertex (lane q writes source q's entry).  Resets the
// counters of buffer (L+1) % 3, which level L+1 fills.
__global__ __launch_bounds__(kBlock) void ms_commit_kernel(MsBufs b, int L) {
    const int nx = L % 3;
    const uint64_t *nxt = b.fr[nx];
    const int32_t *q = b.q[nx];
    const unsigned long long len = b.cnt[nx].len;
    const int lane = lane_id();
    unsigned long long bits = 0;
    const unsigned long long waves = (unsigned long long)gridDim.x * kWavesPerBlock;
    for (unsigned long long i = blockIdx.x * (unsigned long long)kWavesPerBlock +
                                threadIdx.x / kWave;
         i < len; i += waves) {
        const int32_t v = q[i];
        const uint64_t nb = nxt[v];
        if ((nb >> lane) & 1) b.depth[(size_t)v * kMsQ + lane] = L;
        if (lane == 0) {
            b.seen[v] |= nb;
            bits += (unsigned long long)__popcll(nb);
        }
    }
    const unsigned long long t = block_sum_u64(bits);
    if (threadIdx.x == 0 && t) atomicAdd(&b.cnt[nx].bits, t);
    if (blockIdx.x == 0 && threadIdx.x == 0) b.cnt[(L + 1) % 3] = MsCount{0, 0, 0, 0, 0};
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct MsBfsWs {
    void *dmem = nullptr;
    MsCount *h_cnt = nullptr;  // pinned mirror of the three counter slots
    MsBufs b{};
    ~MsBfsWs() {
        if (dmem) hipFree(dmem);
        if (h_cnt) hipHostFree(h_cnt);
    }
};

void free_msbfs_ws(hc_graph_s *g) {
    delete static_cast<MsBfsWs *>(g->ws_msbfs);
    g->ws_msbfs = nullptr;
}

static int get_msbfs_ws(hc_graph_s *g, MsBfsWs **out) {
    auto *ws = static_cast<MsBfsWs *>(g->ws_msbfs);
    if (!ws) {
        ws = new (std::nothrow) MsBfsWs();
        if (!ws) return fail(HC_ERR_OOM, "host allocation failed");
        const size_t n = (size_t)std::max(g->n, 1);
        auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
        const size_t sz_mask = al(sizeof(uint64_t) * n), sz_q = al(sizeof(int32_t) * n);
        const size_t sz_cnt = al(sizeof(MsCount) * 3);
        const size_t total = 4 * sz_mask + 3 * sz_q + sz_cnt;
        char *base = nullptr;
        if (hipMalloc(reinterpret_cast<void **>(&base), total) != hipSuccess) {
            delete ws;
            return fail(HC_ERR_OOM, "batched-BFS workspace of %zu bytes failed", total);
        }
        ws->dmem = base;
        if (hipHostMalloc(reinterpret_cast<void **>(&ws->h_cnt), sizeof(MsCount) * 3,
                          hipHostMallocDefault) != hipSuccess) {
            delete ws;
            return fail(HC_ERR_OOM, "pinned counter allocation failed");
        }
        MsBufs &b = ws->b;
        b.seen = reinterpret_cast<uint64_t *>(base);
        for (int k = 0; k < 3; k++) b.fr[k] = reinterpret_cast<uint64_t *>(base + (k + 1) * sz_mask);
        for (int k = 0; k < 3; k++)
            b.q[k] = reinterpret_cast<int32_t *>(base + 4 * sz_mask + k * sz_q);
        b.cnt = reinterpret_cast<MsCount *>(base + 4 * sz_mask + 3 * sz_q);
        b.n = g->n;
        g->ws_msbfs = ws;
    }
    *out = ws;
    return HC_OK;
}

template <typename OffT>
static int run_msbfs(hc_graph_s *g, const int32_t *sources, int num, int32_t *depth,
                     int max_levels, hc_run_stats_t *stats, hipStream_t s) {
    MsBfsWs *ws = nullptr;
    HC_TRY(get_msbfs_ws(g, &ws));
    const PlanHost *ph = nullptr;
    HC_TRY(get_plan(g, 0, g->n, sssp_plan_config(), &ph));
    // Pull levels gather column block by column block: four blocks on flat
    // graphs (the copy the SSSP sweeps use), sixteen on skewed ones, whose
    // hub rows would otherwise gather from the whole mask array (measured,
    // 64 sources: uniform 23 75 -> 36 ms, Kronecker 24 42 -> 29 ms).
    const ColBlocks *cb = nullptr;
    const PlanHost *pcb = nullptr;
    int shift = tuning().msbfs_col_shift;
    if (shift < 0) {
        int lg = 0;
        while ((int64_t(1) << lg) < (int64_t)g->n) lg++;
        shift = std::max(16, lg - (flat_graph(g) ? 2 : 4));
    }
    if (shift > 0) HC_TRY(get_col_blocks(g, shift, s, &cb));
    if (cb) {
        PlanConfig c = cb_plan_config(cb->K);
        c.min_degree = 1;
        HC_TRY(get_plan(g, 0, g->n, c, &pcb));
    }
    MsBufs b = ws->b;
    b.depth = depth;
    b.full = num >= kMsQ ? ~0ull : ((1ull << num) - 1);
    MsSeeds seeds{};
    for (int q = 0; q < num; q++) {
        int j = 0;
        while (j < seeds.k && seeds.v[j] != sources[q]) j++;
        if (j == seeds.k) {
            seeds.v[j] = sources[q];
            seeds.mask[j] = 0;
            seeds.k++;
        }
        seeds.mask[j] |= 1ull << q;
    }
    const GraphView<OffT> gv = view<OffT>(g);
    const int grid = grid_blocks();
    const int alpha = std::max(1, tuning().msbfs_alpha);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HC_HIP(hipEventCreate(&e0));
    HC_HIP(hipEventCreate(&e1));
    HC_HIP(hipEventRecord(e0, s));
    ms_fill_kernel<<<grid, kBlock, 0, s>>>(b);
    ms_seed_kernel<OffT><<<1, kMsQ, 0, s>>>(b, gv, seeds);
    int rc = hipGetLastError() == hipSuccess ? HC_OK : fail(HC_ERR_HIP, "batched BFS start failed");
    unsigned long long len = (unsigned long long)seeds.k, deg = 0;
    for (int q = 0; q < seeds.k; q++) {
        const int32_t v = seeds.v[q];
        deg += (unsigned long long)(g->h_off[v + 1] - g->h_off[v]);
    }
    const unsigned long long m = (unsigned long long)std::max<int64_t>(g->m, 1);
    const int cap = max_levels > 0 ? max_levels : g->n + 1;
    int levels = 0, pushes = 0, pulls = 0;
    double bytes = 4.0 * kMsQ * (double)g->n + 32.0 * g->n;  // fill
    while (rc == HC_OK && len > 0 && levels < cap) {
        const int L = levels + 1;
        const bool push = deg * (unsigned long long)alpha < m;
        if (push) {
            ms_push_kernel<OffT><<<grid, kBlock, 0, s>>>(b, gv, L);
            pushes++;
        } else if (cb) {
            const GraphView<OffT> cbv{static_cast<const OffT *>(cb->off), cb->adj, nullptr, 0,
                                      g->n};
            ms_pull_cb_kernel<OffT><<<grid, kBlock, 0, s>>>(b, gv, cbv, cb->K, pcb->dev, L);
            pulls++;
        } else {
            ms_pull_kernel<OffT><<<grid, kBlock, 0, s>>>(b, gv, ph->dev, L);
            pulls++;
        }
        ms_commit_kernel<<<grid, kBlock, 0, s>>>(b, L);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(ws->h_cnt + L % 3, b.cnt + L % 3, sizeof(MsCount),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(HC_ERR_HIP, "batched BFS level %d failed", L);
            break;
        }
        const MsCount c = ws->h_cnt[L % 3];
        // Offsets and ids of the examined edges, the 8-byte mask gathered
        // per edge (seen on push, the frontier on pull) and per listed
        // vertex, and the depth entries written.
        bytes += 12.0 * (double)c.edges + 24.0 * (double)c.len + 4.0 * (double)c.bits +
                 (push ? 16.0 * (double)len : 8.0 * g->n);
        len = c.len;
        deg = c.deg;
        levels = L;
    }
    HC_HIP(hipEventRecord(e1, s));
    HC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    HC_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HC_TRY(rc);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = ms;
        stats->epochs = levels;
        stats->push_epochs = pushes;
        stats->pull_epochs = pulls;
        stats->launches = 2 + 2 * levels;
        stats->reached = num;
        stats->alg_bytes = bytes;
    }
    return HC_OK;
}

int run_msbfs_entry(hc_graph_s *g, const int32_t *sources, int num, int32_t *depth,
                    int max_levels, hc_run_stats_t *stats, hipStream_t s) {
    return g->off64 ? run_msbfs<uint64_t>(g, sources, num, depth, max_levels, stats, s)
                    : run_msbfs<uint32_t>(g, sources, num, depth, max_levels, stats, s);
}

}  // namespace hc
