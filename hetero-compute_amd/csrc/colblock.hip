// Column-blocked CSR for pull sweeps, built on the GPU from the resident CSR.
//
// A pull sweep of the plain CSR gathers dist[v] (or contrib[v]) for
// neighbours spread over the whole per-vertex array: on graphs of 2^23
// vertices that array (33.5 MB) misses each XCD's 4 MB L2, and every gather
// waits on the far cache or HBM.  With the edges of each row split by
// neighbour block (block k = ids [k 2^shift, (k+1) 2^shift)), a sweep that
// takes block 0, then block 1, ... gathers from one 2^shift-vertex window at
// a time, which stays in L2 (measured on uniform scale 23: 2.2 ms per sweep
// against 4.4 ms).  Rows keep their CSR order inside each slice.
//
// Build: per (block, row) edge counts -> exclusive scan over the block-major
// [K][n+1] array -> scatter of ids and weights.  Three launch families, once
// per graph (and its shares), a few milliseconds at scale 23.
#include <algorithm>
#include <climits>
#include <cstring>

#include "frontier.h"

namespace hc {

constexpr int kMaxColBlocks = 32;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kBlock * kScanItems;

ColBlockSlot::~ColBlockSlot() {
    for (auto &kv : by_shift)
        if (kv.second.dmem) (void)hipFree(kv.second.dmem);
}

// cnt[k][u] = edges of row u whose neighbour is in block k; cnt[k][n] = 0.
// Per-thread counters in dynamic LDS, K x kBlock words (K blocks only, so
// few blocks leave room for more workgroups per CU).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void cb_count_kernel(GraphView<OffT> g, int shift, int K,
                                                          OffT *cnt) {
    extern __shared__ uint32_t c_dyn[];
    uint32_t(*c)[kBlock] = reinterpret_cast<uint32_t(*)[kBlock]>(c_dyn);
    const int t = threadIdx.x;
    const size_t row_stride = (size_t)g.n + 1;
    const int stride = gridDim.x * kBlock;
    for (int32_t u = blockIdx.x * kBlock + t; u < g.n; u += stride) {
        for (int k = 0; k < K; k++) c[k][t] = 0;
        const OffT b = g.rbeg(u), e = g.rend(u);
        for (OffT i = b; i < e; i++) c[g.adj[i] >> shift][t]++;
        for (int k = 0; k < K; k++) cnt[(size_t)k * row_stride + u] = OffT(c[k][t]);
    }
    const int gid = blockIdx.x * kBlock + t;
    if (gid < K) cnt[(size_t)gid * row_stride + g.n] = 0;
}

// Sum of each kScanChunk-element chunk of x into part.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void cb_scan_reduce_kernel(const OffT *x, size_t len,
                                                                OffT *part) {
    __shared__ OffT s[kBlock];
    const size_t b0 = (size_t)blockIdx.x * kScanChunk;
    const size_t b1 = std::min(len, b0 + kScanChunk);
    OffT v = 0;
    for (size_t i = b0 + threadIdx.x; i < b1; i += kBlock) v += x[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Exclusive scan of the chunk sums in place, by one workgroup: each thread
// sums a contiguous run, the run totals are scanned in LDS, then each thread
// rewrites its run.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void cb_scan_partials_kernel(OffT *part, int nparts) {
    __shared__ OffT s_run[kBlock];
    const int t = threadIdx.x;
    const int per = (nparts + kBlock - 1) / kBlock;
    const int b = std::min(nparts, t * per), e = std::min(nparts, b + per);
    OffT tot = 0;
    for (int i = b; i < e; i++) tot += part[i];
    s_run[t] = tot;
    __syncthreads();
    if (t == 0) {
        OffT run = 0;
        for (int i = 0; i < kBlock; i++) {
            const OffT x = s_run[i];
            s_run[i] = run;
            run += x;
        }
    }
    __syncthreads();
    OffT run = s_run[t];
    for (int i = b; i < e; i++) {
        const OffT v = part[i];
        part[i] = run;
        run += v;
    }
}

// Exclusive scan of each chunk in place, offset by its chunk's prefix.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void cb_scan_apply_kernel(OffT *x, size_t len,
                                                               const OffT *part) {
    __shared__ OffT s_tot[kBlock];
    const size_t base = (size_t)blockIdx.x * kScanChunk + (size_t)threadIdx.x * kScanItems;
    OffT v[kScanItems];
    OffT tot = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        v[j] = base + j < len ? x[base + j] : OffT(0);
        tot += v[j];
    }
    s_tot[threadIdx.x] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
        OffT run = part[blockIdx.x];
        for (int i = 0; i < kBlock; i++) {
            const OffT t = s_tot[i];
            s_tot[i] = run;
            run += t;
        }
    }
    __syncthreads();
    OffT run = s_tot[threadIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        if (base + j < len) x[base + j] = run;
        run += v[j];
    }
}

// Copies each row's edges into their block slices, in CSR order (per-thread
// slice cursors in dynamic LDS, K x kBlock words).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void cb_scatter_kernel(GraphView<OffT> g, int shift, int K,
                                                            const OffT *off, int32_t *adj,
                                                            int32_t *wt) {
    extern __shared__ uint32_t c_dyn[];
    uint32_t(*c)[kBlock] = reinterpret_cast<uint32_t(*)[kBlock]>(c_dyn);
    const int t = threadIdx.x;
    const size_t row_stride = (size_t)g.n + 1;
    const int stride = gridDim.x * kBlock;
    for (int32_t u = blockIdx.x * kBlock + t; u < g.n; u += stride) {
        for (int k = 0; k < K; k++) c[k][t] = 0;
        const OffT b = g.rbeg(u), e = g.rend(u);
        for (OffT i = b; i < e; i++) {
            const int32_t v = g.adj[i];
            const int k = v >> shift;
            const OffT p = off[(size_t)k * row_stride + u] + OffT(c[k][t]++);
            adj[p] = v;
            if (wt) wt[p] = g.wt[i];
        }
    }
}

template <typename OffT>
static int build_col_blocks(hc_graph_s *g, int shift, int K, hipStream_t s, ColBlocks *cb) {
    const size_t n = (size_t)g->n, m = (size_t)g->m;
    const size_t len = (size_t)K * (n + 1);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t sz_off = al(sizeof(OffT) * len), sz_adj = al(sizeof(int32_t) * std::max<size_t>(m, 1));
    const size_t sz_wt = g->weighted ? sz_adj : 0;
    const size_t sz_seen = al(sizeof(uint32_t) * ((n + 31) / 32));
    const size_t total = sz_off + sz_adj + sz_wt + sz_seen;
    char *base = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&base), total) != hipSuccess)
        return fail(HC_ERR_OOM, "column-blocked copy of %zu bytes failed", total);
    const int nparts = (int)((len + kScanChunk - 1) / kScanChunk);
    OffT *part = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&part), sizeof(OffT) * nparts) != hipSuccess) {
        (void)hipFree(base);
        return fail(HC_ERR_OOM, "column-block scan buffer failed");
    }
    OffT *off = reinterpret_cast<OffT *>(base);
    int32_t *adj = reinterpret_cast<int32_t *>(base + sz_off);
    int32_t *wt = g->weighted ? reinterpret_cast<int32_t *>(base + sz_off + sz_adj) : nullptr;
    uint32_t *seen = reinterpret_cast<uint32_t *>(base + sz_off + sz_adj + sz_wt);
    const GraphView<OffT> v = view<OffT>(g);
    const int rgrid = std::max(1, std::min(grid_blocks(), (int)((n + kBlock - 1) / kBlock)));
    const size_t lds = sizeof(uint32_t) * (size_t)K * kBlock;
    cb_count_kernel<OffT><<<std::max(rgrid, (K + kBlock - 1) / kBlock), kBlock, lds, s>>>(
        v, shift, K, off);
    cb_scan_reduce_kernel<OffT><<<nparts, kBlock, 0, s>>>(off, len, part);
    cb_scan_partials_kernel<OffT><<<1, kBlock, 0, s>>>(part, nparts);
    cb_scan_apply_kernel<OffT><<<nparts, kBlock, 0, s>>>(off, len, part);
    cb_scatter_kernel<OffT><<<rgrid, kBlock, lds, s>>>(v, shift, K, off, adj, wt);
    int rc = HC_OK;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        rc = fail(HC_ERR_HIP, "column-block build failed");
    (void)hipFree(part);
    if (rc != HC_OK) {
        (void)hipFree(base);
        return rc;
    }
    cb->shift = shift;
    cb->K = K;
    cb->off = off;
    cb->adj = adj;
    cb->wt = wt;
    cb->seen = seen;
    cb->dmem = base;
    cb->bytes = total;
    return HC_OK;
}

int col_shift_for(const hc_graph_s *g, bool flat) {
    const int knob = tuning().pull_col_shift;
    if (knob == 0 || g->row0 != 0 || g->rows != g->n || g->n <= 0) return 0;
    int sh = knob;
    if (knob < 0) {
        // Automatic: flat graphs only (on skewed ones the hubs' rows already
        // gather from the few L2-resident hub entries, and slicing them by
        // block measured slower), four blocks of at least 2^20 vertices
        // (measured on uniform scale 23: four blocks beat eight, whose
        // thinner slices cost more per row than their smaller window saves).
        if (!flat) return 0;
        int lg = 0;
        while ((int64_t(1) << lg) < (int64_t)g->n) lg++;
        sh = std::max(20, lg - 2);
    }
    if ((((int64_t)g->n + (int64_t(1) << sh) - 1) >> sh) < 2) return 0;
    // At most kMaxColBlocks blocks: larger blocks for very large graphs.
    while ((((int64_t)g->n + (int64_t(1) << sh) - 1) >> sh) > kMaxColBlocks) sh++;
    return sh;
}

int get_col_blocks(hc_graph_s *g, int shift, hipStream_t s, const ColBlocks **out) {
    *out = nullptr;
    if (shift <= 0) return HC_OK;
    const int K = (int)(((int64_t)g->n + (int64_t(1) << shift) - 1) >> shift);
    if (K < 2 || K > kMaxColBlocks || g->row0 != 0 || g->rows != g->n) return HC_OK;
    ColBlockSlot &slot = *g->cbslot;
    std::lock_guard<std::mutex> lk(slot.mu);
    auto it = slot.by_shift.find(shift);
    if (it != slot.by_shift.end()) {
        *out = &it->second;  // map nodes stay put as others are added
        return HC_OK;
    }
    ColBlocks cb;
    HC_TRY(g->off64 ? build_col_blocks<uint64_t>(g, shift, K, s, &cb)
                    : build_col_blocks<uint32_t>(g, shift, K, s, &cb));
    *out = &slot.by_shift.emplace(shift, cb).first->second;
    return HC_OK;
}

}  // namespace hc
