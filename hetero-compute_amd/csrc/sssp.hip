// SSSP epoch kernels for gfx950.
//
//  * Pull epoch (Bellman-Ford, in place): dist[u] = min(dist[u],
//    min_{(v,w)} dist[v]+w), the arithmetic of epoch_sssp_pull_cpu_one_to_one
//    (src/kernels/cpu/sssp.cuh:155-182) with int32 weights and a saturating
//    sentinel.  Vertex mappings: thread / wavefront / workgroup per vertex
//    (counterparts of src/kernels/gpu/sssp.cuh:120-281) and the fused
//    degree-bucketed plan (hub chunks with atomicMin, workgroup, wavefront,
//    16-lane, 4-lane and 1-lane groups in one launch).
//  * Push epoch: relax the out-edges of the active frontier with atomicMin and
//    append improved vertices (deduplicated through a bitmap) to the next
//    frontier, wavefront-aggregated.
//  * Device-driven loop (step / control / compact launches per epoch, see
//    frontier.h): the controller picks the next direction (pull when the
//    frontier's degree sum exceeds m/alpha) and stops the loop when nothing
//    changed, so the host only enqueues launches.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>
#include <cstring>

#include "frontier.h"

namespace hc {

// ---------------------------------------------------------------------------
// Row reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool changed_bit(const Staged<uint32_t> &bm, int32_t v) {
    return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Minimum candidate over edges [b, e) taken G apart starting at b+lane.
// Trips of four edges per lane with predicated loads (a short row is one
// trip of parallel loads, not a serial tail), and the next trip's ids and
// weights are loaded before this trip's distance gathers, so each trip costs
// about one memory round trip instead of two.
// FILTER: only neighbours set in `changed` (the vertices whose distance
// changed in the previous epoch) are gathered; the others cannot lower the
// row below its current value.  `gathered` counts the gathers.
// Neighbour distances through a 16-bit mirror: half the bytes per gathered
// entry, so twice as many vertices fit each cache.  0xFFFF: unreached;
// 0xFFFE: too large for 16 bits, read the int32 array.  The mirror is
// rebuilt from the int32 distances before every pull epoch and written
// during the sweep by whoever lowers an entry; a row lowered concurrently by
// several hub chunks may keep a stale (larger) mirror entry until the next
// rebuild, which costs nothing in exactness: the row is marked changed, so
// another epoch runs, and it starts from a fresh mirror.
struct Dist16Reader {
    const uint16_t *h;
    const int32_t *full;
    __device__ __forceinline__ int32_t operator[](int32_t v) const {
        const uint32_t x = h[v];
        return x < 0xFFFEu ? int32_t(x) : (x == 0xFFFFu ? HC_INF_DIST : full[v]);
    }
};

__device__ __forceinline__ uint16_t dist16(int32_t d) {
    return d >= HC_INF_DIST ? uint16_t(0xFFFF) : (d >= 0xFFFE ? uint16_t(0xFFFE) : uint16_t(d));
}

template <int G, bool FILTER, typename DR, typename OffT>
__device__ __forceinline__ int32_t row_min(const GraphView<OffT> &g,
                                           const DR &dist,
                                           const Staged<uint32_t> &changed,
                                           OffT b, OffT e, int lane, uint32_t &gathered) {
    int32_t m = HC_INF_DIST;
    OffT i = b + OffT(lane);
    if (i >= e) return m;
    int32_t v[4], w[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const OffT k = i + OffT(r * G);
        v[r] = k < e ? ldnt(&g.adj[k]) : -1;
        w[r] = k < e ? ldnt(&g.wt[k]) : 0;
    }
    for (;;) {
        const OffT ni = i + OffT(4 * G);
        int32_t nv[4] = {-1, -1, -1, -1}, nw[4] = {0, 0, 0, 0};
        if (ni < e) {  // short rows skip the prefetch instructions entirely
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const OffT k = ni + OffT(r * G);
                nv[r] = k < e ? ldnt(&g.adj[k]) : -1;
                nw[r] = k < e ? ldnt(&g.wt[k]) : 0;
            }
        }
        bool use[4];
#pragma unroll
        for (int r = 0; r < 4; r++) use[r] = v[r] >= 0 && (!FILTER || changed_bit(changed, v[r]));
        int32_t d[4];
#pragma unroll
        for (int r = 0; r < 4; r++) d[r] = use[r] ? dist[v[r]] : HC_INF_DIST;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            m = min(m, relax(d[r], w[r]));
            gathered += use[r] ? 1u : 0u;
        }
        if (ni >= e) break;
        for (int r = 0; r < 4; r++) {
            v[r] = nv[r];
            w[r] = nw[r];
        }
        i = ni;
    }
    return m;
}

// Workgroup minimum for any blockDim that is a multiple of 64 (<= 1024).
// The leading barrier keeps the LDS slots from being overwritten while a
// previous call is still reading them (the reuse race of
// src/kernels/gpu/sssp.cuh:259-275).
__device__ __forceinline__ int32_t block_min(int32_t m) {
    __shared__ int32_t s_min[16];
    m = group_min<64>(m);
    __syncthreads();
    if (lane_id() == 0) s_min[threadIdx.x / kWave] = m;
    __syncthreads();
    int32_t r = s_min[0];
    const int nw = blockDim.x / kWave;
    for (int i = 1; i < nw; i++) r = min(r, s_min[i]);
    return r;
}

// ---------------------------------------------------------------------------
// Fused degree-bucketed pull sweep
// ---------------------------------------------------------------------------

// Lowers dist[u] to m when that is lower; chg: this call lowered it.  For
// rows that several groups of one sweep may lower (hub chunks, the column
// blocks of a row): atomicMin decides, a stale `cur` only costs an atomic.
__device__ __forceinline__ void lower_shared_row(int32_t *dist, uint16_t *d16, int32_t u,
                                                 int32_t cur, int32_t m, bool &chg) {
    if (m < cur) {
        const int32_t old = atomicMin(&dist[u], m);
        chg = m < old;
        // Two groups lowering one row may store their mirror values out of
        // order; a stale (larger) entry only weakens this epoch, and the row
        // is marked changed, so the next epoch starts from a rebuild.
        if (chg && d16) d16[u] = dist16(m);
    }
}

// One row of class c per group of G lanes.  SHARED: other groups of the
// sweep may lower the same row (column-blocked sweeps), so the row is
// lowered with atomicMin.
template <int G, bool FILTER, bool SHARED = false, typename DR, typename OffT>
__device__ __forceinline__ void pull_group(const PlanDev &p, int c, int k,
                                           const GraphView<OffT> &g,
                                           const DR &rd,
                                           const Staged<uint32_t> &changed,
                                           int32_t *__restrict__ dist, uint16_t *d16,
                                           bool &chg, int32_t &u, uint32_t &gathered,
                                           int64_t &deg) {
    constexpr int PER = kBlock / G;
    const int gid = threadIdx.x / G, lane = threadIdx.x % G;
    const int idx = k * PER + gid;
    if (idx < p.cnt[c]) {
        int32_t uu = p.list[c] ? p.list[c][idx] : p.beg[c] + idx;
        const OffT b = g.rbeg(uu), e = g.rend(uu);
        // The row's own distance is read with its bounds, before the
        // gathers: in a pull epoch only this group writes it (unless SHARED).
        const int32_t cur = lane == 0 ? dist[uu] : 0;
        int32_t m = row_min<G, FILTER>(g, rd, changed, b, e, lane, gathered);
        if (G > 1) m = group_min<G>(m);
        if (lane == 0) {
            if (SHARED) {
                lower_shared_row(dist, d16, uu, cur, m, chg);
            } else if (m < cur) {
                dist[uu] = m;
                if (d16) d16[uu] = dist16(m);
                chg = true;
            }
            u = uu;
            deg = int64_t(e - b);
        }
    }
}


// One pull sweep over the plan's units (grid-stride).  LOOP: changed
// vertices are set in the next-frontier bitmap and counted (first setting
// only: hub chunks may lower the same vertex twice); otherwise changes are
// counted for the epoch API.
// Neighbour distances are read through `rd` (hub entries from the LDS
// snapshot); the row's own distance is read and written in global memory.
// A stale snapshot only delays an improvement to a later epoch: values only
// decrease, and an epoch without any change had nothing stale to read.
// FILTER: gather only neighbours in `changed` (see row_min); `gathered`
// counts the gathers.
template <bool LOOP, bool FILTER, typename DR, typename OffT>
__device__ void sssp_pull_sweep(const PlanDev &p, const GraphView<OffT> &g,
                                const DR &rd, const Staged<uint32_t> &changed,
                                int32_t *__restrict__ dist, uint16_t *d16, uint32_t *next_bm,
                                uint32_t &cnt, unsigned long long &edges_acc,
                                uint32_t &gathered) {
    const int total = p.unit_start[kNumClasses];
    for (int unit = blockIdx.x; unit < total; unit += gridDim.x) {
        int c = 0;
        while (unit >= p.unit_start[c + 1]) c++;
        const int k = unit - p.unit_start[c];
        bool chg = false;
        int32_t u = 0;
        int64_t deg = -1;  // the row's degree when its bounds were read here
        if (c <= 1) {
            int32_t uu;
            OffT b, e;
            if (c == 0) {
                HubChunk h = p.chunks[k];
                uu = h.u;
                b = OffT(h.b);
                e = OffT(h.e);
            } else {
                uu = p.list[1] ? p.list[1][k] : p.beg[1] + k;
                b = g.rbeg(uu);
                e = g.rend(uu);
                deg = int64_t(e - b);
            }
            int32_t m = block_min(row_min<kBlock, FILTER>(g, rd, changed, b, e, threadIdx.x,
                                                          gathered));
            if (threadIdx.x == 0) {
                if (c == 0) {
                    lower_shared_row(dist, d16, uu, dist[uu], m, chg);
                } else {
                    int32_t cur = dist[uu];
                    if (m < cur) {
                        dist[uu] = m;
                        if (d16) d16[uu] = dist16(m);
                        chg = true;
                    }
                }
                u = uu;
            }
        } else if (c == 2) {
            pull_group<64, FILTER>(p, c, k, g, rd, changed, dist, d16, chg, u, gathered, deg);
        } else if (c == 3) {
            pull_group<16, FILTER>(p, c, k, g, rd, changed, dist, d16, chg, u, gathered, deg);
        } else if (c == 4) {
            pull_group<4, FILTER>(p, c, k, g, rd, changed, dist, d16, chg, u, gathered, deg);
        } else {
            pull_group<1, FILTER>(p, c, k, g, rd, changed, dist, d16, chg, u, gathered, deg);
        }
        if (LOOP && c == 0) {
            // Several chunks may lower the same hub: count its bit once.
            if (wave_set_bit(next_bm, u, chg)) {
                cnt++;
                edges_acc += (unsigned long long)(g.rend(u) - g.rbeg(u));
            }
        } else if (LOOP) {
            // Any other row is swept by one group only.
            wave_or_bits(next_bm, u, chg);
            if (chg) {
                cnt++;
                edges_acc += (unsigned long long)(deg >= 0 ? deg : int64_t(g.rend(u) - g.rbeg(u)));
            }
        } else {
            cnt += chg ? 1u : 0u;
        }
    }
}

// One pull sweep over the column blocks of the graph (see colblock.hip):
// units (block k, plan unit) in block-major order, so the grid gathers from
// one block of distances at a time.  `cb` views the blocked arrays with the
// offsets of block 0; block k's slices start (n+1) k offsets later.  Plan
// classes are by mean slice degree and have no hub chunks.  A row may be
// lowered by several blocks in one sweep: atomicMin, and `counted` (the
// next-frontier bitmap, or the epoch API's seen bitmap) counts it once;
// edges_acc adds its full degree.
template <typename DR, typename OffT>
__device__ void sssp_pull_sweep_cb(const PlanDev &p, const GraphView<OffT> &g,
                                   const GraphView<OffT> &cb, int K, const DR &rd,
                                   int32_t *__restrict__ dist, uint16_t *d16,
                                   uint32_t *counted, uint32_t &cnt,
                                   unsigned long long &edges_acc, uint32_t &gathered) {
    const int U = p.unit_start[kNumClasses];
    const int total = U * K;
    const Staged<uint32_t> nochg = unstaged<uint32_t>(nullptr);
    for (int unit = blockIdx.x; unit < total; unit += gridDim.x) {
        const int k = unit / U;
        const int pu = unit - k * U;
        GraphView<OffT> gk = cb;
        gk.off = cb.off + (size_t)k * ((size_t)g.n + 1);
        int c = 1;
        while (pu >= p.unit_start[c + 1]) c++;
        const int kk = pu - p.unit_start[c];
        bool chg = false;
        int32_t u = 0;
        int64_t deg = -1;
        if (c == 1) {
            const int32_t uu = p.list[1] ? p.list[1][kk] : p.beg[1] + kk;
            const int32_t m = block_min(row_min<kBlock, false>(gk, rd, nochg, gk.rbeg(uu),
                                                               gk.rend(uu), threadIdx.x,
                                                               gathered));
            if (threadIdx.x == 0) {
                lower_shared_row(dist, d16, uu, dist[uu], m, chg);
                u = uu;
            }
        } else if (c == 2) {
            pull_group<64, false, true>(p, c, kk, gk, rd, nochg, dist, d16, chg, u, gathered, deg);
        } else if (c == 3) {
            pull_group<16, false, true>(p, c, kk, gk, rd, nochg, dist, d16, chg, u, gathered, deg);
        } else if (c == 4) {
            pull_group<4, false, true>(p, c, kk, gk, rd, nochg, dist, d16, chg, u, gathered, deg);
        } else {
            pull_group<1, false, true>(p, c, kk, gk, rd, nochg, dist, d16, chg, u, gathered, deg);
        }
        if (wave_set_bit(counted, u, chg)) {
            cnt++;
            edges_acc += (unsigned long long)(g.rend(u) - g.rbeg(u));
        }
    }
}

// ---------------------------------------------------------------------------
// Push expansion of the frontier
// ---------------------------------------------------------------------------

// Relaxes the out-edges [b, e) of u (G lanes); a lowered neighbour joins
// the next frontier, or, when its new value is at or above `thresh`
// (buckets), the far bitmap (`nfar` counts those settings).
template <int G, typename OffT>
__device__ __forceinline__ void push_range(const GraphView<OffT> &g,
                                           int32_t *__restrict__ dist,
                                           const FrontierBufs &fb, StepCtl *ctl,
                                           const PushOut &out,
                                           int nxt, uint32_t *wbm, int32_t u, OffT b,
                                           OffT e, int lane, uint32_t &cnt,
                                           unsigned long long &edges_acc, int32_t thresh,
                                           uint32_t &nfar) {
    const int32_t du = dist[u];
    for (OffT i = b + OffT(lane); i < e; i += OffT(G)) {
        int32_t v = g.adj[i];
        int32_t nd = relax(du, g.wt[i]);
        bool lowered = false;
        if (nd < dist[v]) lowered = nd < atomicMin(&dist[v], nd);
        if (lowered && nd >= thresh) {
            atomicOr(&fb.far[v >> 5], 1u << (v & 31));
            nfar++;
            lowered = false;
        }
        if (!out.dedup && lowered) {
            // The first edge to set v's bit appends v, now.
            const uint32_t bit = 1u << (v & 31);
            lowered = !(atomicOr(&wbm[v >> 5], bit) & bit);
            if (lowered) {
                cnt++;
                edges_acc += (unsigned long long)(g.rend(v) - g.rbeg(v));
            }
        }
        // Deferred: a lowered neighbour is a candidate; the staging keeps
        // its first copy (bit `wbm`) and counts it there.
        push_out_append(out, g, fb, ctl, nxt, lowered, v, cnt, edges_acc);
    }
}

template <typename OffT>
__device__ void sssp_push_sweep(const GraphView<OffT> &g,
                                int32_t *__restrict__ dist,
                                const FrontierBufs &fb, StepCtl *ctl,
                                const PushOut &out, int cur,
                                int nxt, uint32_t *wbm, uint32_t &cnt,
                                unsigned long long &edges_acc, uint32_t &nfar) {
    const int32_t thresh = ctl->thresh;
    const uint32_t nch = fb.cnt[cur][3], n0 = fb.cnt[cur][0],
                   n1 = fb.cnt[cur][1], n2 = fb.cnt[cur][2];
    const uint32_t u1 = nch + n0;
    const uint32_t u2 = u1 + (n1 + kWavesPerBlock - 1) / kWavesPerBlock;
    const uint32_t total = u2 + (n2 + kSmallPerUnit - 1) / kSmallPerUnit;
    // An upper bound of the staged count: the workgroup synchronizes to check
    // for room only when the next unit could overfill the buffer.
    uint32_t staged = 0;
    for (uint32_t unit = blockIdx.x; unit < total; unit += gridDim.x) {
        const uint32_t umax = push_unit_max(unit, nch, ctl->chunk);
        if (staged + umax > out.cap)
            staged = push_out_make_room(out, g, fb, ctl, nxt, umax, cnt, edges_acc);
        staged += umax;
        if (unit < nch) {
            HubChunk h = fb.ch[cur][unit];
            push_range<kBlock>(g, dist, fb, ctl, out, nxt, wbm, h.u, OffT(h.b), OffT(h.e),
                               threadIdx.x, cnt, edges_acc, thresh, nfar);
        } else if (unit < u1) {
            int32_t u = fb.q[cur][0][unit - nch];
            push_range<kBlock>(g, dist, fb, ctl, out, nxt, wbm, u, g.rbeg(u), g.rend(u),
                               threadIdx.x, cnt, edges_acc, thresh, nfar);
        } else if (unit < u2) {
            uint32_t idx = (unit - u1) * kWavesPerBlock + threadIdx.x / kWave;
            if (idx < n1) {
                int32_t u = fb.q[cur][1][idx];
                push_range<kWave>(g, dist, fb, ctl, out, nxt, wbm, u, g.rbeg(u), g.rend(u),
                                  lane_id(), cnt, edges_acc, thresh, nfar);
            }
        } else {
            uint32_t idx = (unit - u2) * kSmallPerUnit + threadIdx.x / kSmallLanes;
            if (idx < n2) {
                int32_t u = fb.q[cur][2][idx];
                push_range<kSmallLanes>(g, dist, fb, ctl, out, nxt, wbm, u, g.rbeg(u),
                                        g.rend(u), threadIdx.x % kSmallLanes, cnt,
                                        edges_acc, thresh, nfar);
            }
        }
    }
}

// Promotion epoch of the buckets: every far vertex whose distance is now
// below `thresh` leaves the far bitmap and is set in the changed bitmap
// `wbm` (compacted into the next frontier's queues by the launch after);
// `cnt` / `edges_acc` count those, `left` the vertices still far.  A thread
// owns one word of both bitmaps.
template <typename OffT>
__device__ void sssp_promote_sweep(const GraphView<OffT> &g, const int32_t *dist,
                                   const FrontierBufs &fb, int32_t thresh, uint32_t *wbm,
                                   uint32_t &cnt, unsigned long long &edges_acc, uint32_t &left) {
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < fb.words; w += gridDim.x * kBlock) {
        const uint32_t word = fb.far[w];
        if (!word) continue;
        uint32_t now = 0;
        for (uint32_t x = word; x; x &= x - 1) {
            const int b = __ffs(x) - 1;
            const int32_t v = w * 32 + b;
            if (dist[v] < thresh) {
                now |= 1u << b;
                cnt++;
                edges_acc += (unsigned long long)(g.rend(v) - g.rbeg(v));
            } else {
                left++;
            }
        }
        if (now) {
            fb.far[w] = word & ~now;
            wbm[w] = now;  // cleared by the launch before; no other writer
        }
    }
}

// ---------------------------------------------------------------------------
// Kernels of the device-driven loop
// ---------------------------------------------------------------------------

// `nstage` hub distances are staged in dynamic LDS for pull epochs; push
// epochs stage up to `push_cap` output entries there.  D16: pull
// gathers read the 16-bit mirror `d16` instead.
template <typename OffT, bool D16, bool CB>
__global__ __launch_bounds__(kBlock) void sssp_step_kernel(
    PlanDev p, GraphView<OffT> g, int32_t *__restrict__ dist, uint16_t *d16, FrontierBufs fb,
    StepCtl *ctl, int32_t nstage, int32_t push_cap, int defer_dedup, PlanDev pcb,
    GraphView<OffT> cbv, int cbk) {
    extern __shared__ int32_t s_stage[];
    if (ctl->done) return;
    const int mode = ctl->mode;
    const int k = ctl->epoch;
    const int cur = k & 1, nxt = cur ^ 1;  // queue parities
    // Bitmaps: changed in epoch k-1 (read), changed in epoch k (written),
    // and the one to clear for epoch k+1.
    const uint32_t *rbm = fb.bm[k % 3];
    uint32_t *wbm = fb.bm[(k + 1) % 3];
    uint32_t *cbm = fb.bm[(k + 2) % 3];
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < fb.words; w += gridDim.x * kBlock)
        cbm[w] = 0;
    uint32_t cnt = 0, gathered = 0;
    unsigned long long eacc = 0;
    if (mode == kPull && CB) {
        // Column-blocked sweep (the host picks it only when filtering is
        // not forced; its sweeps gather every neighbour):
        // the heavy rows of pcb block by block, then the light rows of plan
        // p once over the plain CSR (heavy rows first, as in the plain
        // sweep's class order, so light rows read their neighbours' values
        // of this epoch).
        if (D16) {
            const Dist16Reader rd{d16, dist};
            sssp_pull_sweep_cb(pcb, g, cbv, cbk, rd, dist, d16, wbm, cnt, eacc, gathered);
            sssp_pull_sweep<true, false>(p, g, rd, unstaged(rbm), dist, d16, wbm, cnt, eacc,
                                         gathered);
        } else {
            const Staged<int32_t> rd = stage(dist, nstage, s_stage);
            sssp_pull_sweep_cb(pcb, g, cbv, cbk, rd, dist, nullptr, wbm, cnt, eacc, gathered);
            sssp_pull_sweep<true, false>(p, g, rd, unstaged(rbm), dist, nullptr, wbm, cnt, eacc,
                                         gathered);
        }
    } else if (mode == kPull) {
        // One gather path per instantiation: both in one kernel would make
        // it hold the registers of both (and spill).
        if (D16) {
            const Dist16Reader rd{d16, dist};
            if (ctl->filter)
                sssp_pull_sweep<true, true>(p, g, rd, unstaged(rbm), dist, d16, wbm, cnt, eacc,
                                            gathered);
            else
                sssp_pull_sweep<true, false>(p, g, rd, unstaged(rbm), dist, d16, wbm, cnt, eacc,
                                             gathered);
        } else {
            const Staged<int32_t> rd = stage(dist, nstage, s_stage);
            if (ctl->filter)
                sssp_pull_sweep<true, true>(p, g, rd, unstaged(rbm), dist, nullptr, wbm, cnt,
                                            eacc, gathered);
            else
                sssp_pull_sweep<true, false>(p, g, rd, unstaged(rbm), dist, nullptr, wbm, cnt,
                                             eacc, gathered);
        }
    } else if (mode == kPromote) {
        // (`gathered` carries the vertices left far)
        sssp_promote_sweep(g, dist, fb, ctl->thresh, wbm, cnt, eacc, gathered);
    } else {
        // Output staged in the dynamic LDS (unused by push epochs).
        __shared__ uint32_t s_out;
        if (threadIdx.x == 0) s_out = 0;
        __syncthreads();
        const PushOut out{s_stage, &s_out, (uint32_t)push_cap, defer_dedup ? wbm : nullptr};
        // (`gathered` carries the far settings)
        sssp_push_sweep(g, dist, fb, ctl, out, cur, nxt, wbm, cnt, eacc, gathered);
        const PushKept k = push_out_flush(out, g, fb, ctl, nxt);
        cnt += k.n;
        eacc += k.e;
    }
    write_partials(fb.part, cnt, eacc, gathered);
}

// Epoch controller (one thread, on the header copy c), from the step sums
// of the epoch that ran.
__device__ void sssp_decide(StepCtl *c, StepCtl *trace,
                            const unsigned long long (&sums)[kSumSlots]) {
    const unsigned long long out_n = sums[0], out_e = sums[1], gathered = sums[2];
    const int k = c->epoch;
    const int ran = c->mode;
    if (ran == kPromote) {
        // far bitmap words, plus dist[v] and a bitmap entry per far vertex
        c->alg_bytes += 4ull * (unsigned long long)((c->n + 31) / 32) + 8ull * (out_n + gathered);
        c->promotions++;
        c->far_live = (long long)gathered;
        trace_epoch(c, trace, k, ran, out_n, out_e, 0);
    } else if (ran == kPush) {
        // offsets + dist[u] per frontier vertex; id + weight + dist[v] per
        // edge; dist write + bitmap/queue entry per improved vertex.
        c->alg_bytes += 12ull * c->in_n + 12ull * c->in_e + 8ull * out_n;
        c->push_epochs++;
        c->far_live += (long long)gathered;  // far settings (may repeat a vertex)
        trace_epoch(c, trace, k, ran, out_n, out_e, c->in_e);
    } else {
        // Offsets, ids and own distances always; weight + distance per
        // gathered edge; one bit per tested neighbour when filtering.
        c->alg_bytes += (unsigned long long)c->sweep_bytes - 8ull * (unsigned long long)c->m +
                        8ull * gathered + (c->filter ? (unsigned long long)c->m / 8 : 0ull) +
                        8ull * out_n;
        c->pull_epochs++;
        trace_epoch(c, trace, k, ran, out_n, out_e, (unsigned long long)c->m);
    }
    c->epoch = k + 1;
    c->in_n = (long long)out_n;
    c->in_e = (long long)out_e;
    // The changed set of this epoch is complete: the next pull may gather
    // only from it (knob 1: always; k >= 2: when it holds under n / k
    // vertices, so that most gathers are skipped).
    c->filter = c->filter_knob == 1 ||
                (c->filter_knob >= 2 && (long long)out_n * c->filter_knob < (long long)c->n);
    int next = ran;
    if (c->max_epochs > 0 && k + 1 >= c->max_epochs) {
        c->done = 1;
    } else if (out_n == 0) {
        if (c->delta > 0 && c->far_live > 0) {
            // Raise the threshold and promote; after a promotion that found
            // nothing below it, rise twice as far.
            const long long step = ran == kPromote ? 2ll * c->far_step : (long long)c->delta;
            c->far_step = (int32_t)min(step, (long long)HC_INF_DIST);
            c->thresh = (int32_t)min((long long)c->thresh + step, (long long)HC_INF_DIST);
            next = kPromote;
        } else {
            c->done = 1;
        }
    } else if (c->delta > 0 || c->direction == 2) {
        next = kPush;
    } else if (c->direction == 1) {
        next = kPull;
    } else {
        next = (long long)out_e > c->m / c->alpha ? kPull : kPush;
    }
    c->need_compact = (!c->done && ran != kPush && next == kPush) ? 1 : 0;
    c->mode = next;
}

// Between epochs k and k + 1: the controller (every workgroup decides on its
// own copy; workgroup 0 records the epoch and stores header k + 1), then
// the frontier compaction when the loop turns to push and, with the 16-bit
// mirror, its rebuild before a pull epoch.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void sssp_between_kernel(GraphView<OffT> g, FrontierBufs fb,
                                                              StepCtl *cur, StepCtl *nxt,
                                                              StepCtl *trace, Partials next_set,
                                                              const int32_t *dist, uint16_t *d16) {
    __shared__ StepCtl s_c;  // header fields only
    unsigned long long sums[kSumSlots];
    if (!ctl_begin(&s_c, cur, fb.part, sums)) return;
    if (threadIdx.x == 0) sssp_decide(&s_c, blockIdx.x == 0 ? trace : nullptr, sums);
    __syncthreads();
    ctl_end(&s_c, cur, nxt, fb, next_set);
    compact_body(g, fb, &s_c, 1, dist, d16);
}

// Initialises distances, bitmaps and the control block for one run.
// init_dist == nullptr: single source `src`; otherwise copy init_dist and
// start in pull mode with an empty frontier (sssp_pull_gpu semantics).
template <typename OffT>
__global__ __launch_bounds__(kBlock) void sssp_init_kernel(
    GraphView<OffT> g, int32_t *__restrict__ dist,
    const int32_t *__restrict__ init_dist, FrontierBufs fb, StepCtl *ctl,
    int32_t src, int direction, int alpha, int max_epochs, long long m,
    long long sweep_bytes, int grid, int filter, int chunk, uint16_t *d16, int delta) {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < g.n; i += stride) {
        const int32_t d = init_dist ? init_dist[i] : (i == src ? 0 : HC_INF_DIST);
        dist[i] = d;
        if (d16) d16[i] = dist16(d);
    }
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < fb.words; w += stride) {
        // Epoch 0 reads bitmap 0 as the set changed before it: the source.
        fb.bm[0][w] = (!init_dist && w == (src >> 5)) ? (1u << (src & 31)) : 0u;
        fb.bm[1][w] = 0;
        fb.bm[2][w] = 0;
        fb.far[w] = 0;
    }
    if (blockIdx.x == 0)
        for (int j = 0; j < kSumSets; j++) clear_partials(sum_set(fb.part, j));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // With initial distances the changed set before epoch 0 is unknown:
        // that epoch gathers every neighbour.
        ctl->filter_knob = filter;
        ctl->filter = init_dist ? 0 : (filter == 1 || (filter >= 2 && filter < g.n));
        for (int p = 0; p < 2; p++)
            for (int j = 0; j < 5; j++) fb.cnt[p][j] = 0;
        ctl->done = 0;
        ctl->epoch = 0;
        ctl->max_epochs = max_epochs;
        ctl->direction = direction;
        ctl->alpha = alpha;
        ctl->need_compact = 0;
        ctl->chunk = chunk;
        ctl->n = g.n;
        ctl->grid = grid;
        ctl->m = m;
        ctl->sweep_bytes = sweep_bytes;
        ctl->push_epochs = ctl->pull_epochs = 0;
        ctl->delta = delta;
        ctl->thresh = delta > 0 ? delta : HC_INF_DIST;
        ctl->far_step = delta;
        ctl->promotions = 0;
        ctl->far_live = 0;
        ctl->reached = 0;
        ctl->alg_bytes = 0;
        ctl->t_start = wall_clock64();
        ctl->t_end = ctl->t_start;
        if (init_dist) {
            ctl->mode = kPull;
            ctl->in_n = g.n;
            ctl->in_e = m;
        } else {
            OffT b = g.rbeg(src), e = g.rend(src);
            int64_t d = int64_t(e - b);
            int cls = degree_class(d);
            if (cls == 3) {  // chunk descriptors: written by workgroup 0 below
                fb.cnt[0][3] = (uint32_t)((d + chunk - 1) / chunk);
                fb.cnt[0][4] = 1;
            } else {
                fb.q[0][cls][0] = src;
                fb.cnt[0][cls] = 1;
            }
            ctl->in_n = 1;
            ctl->in_e = d;
            ctl->mode = direction == 1 ? kPull
                        : (direction == 2 || delta > 0) ? kPush
                        : (d > m / alpha ? kPull : kPush);
        }
    }
    if (blockIdx.x == 0 && !init_dist) write_source_chunks(g, fb, src, chunk);
}

template <typename OffT>
__global__ void sssp_pull_thread_kernel(GraphView<OffT> g,
                                        int32_t *__restrict__ dist,
                                        int32_t start, int32_t end,
                                        int32_t *updated) {
    uint32_t upd = 0;
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t u = start + blockIdx.x * blockDim.x + threadIdx.x; u < end;
         u += stride) {
        uint32_t gathered = 0;
        int32_t m = row_min<1, false>(g, unstaged(dist), unstaged<uint32_t>(nullptr), g.rbeg(u),
                                      g.rend(u), 0, gathered);
        if (m < dist[u]) {
            dist[u] = m;
            upd++;
        }
    }
    for (int s = 32; s > 0; s >>= 1) upd += __shfl_xor(upd, s, 64);
    if (lane_id() == 0 && upd) atomicAdd(updated, (int32_t)upd);
}

template <typename OffT>
__global__ void sssp_pull_wave_kernel(GraphView<OffT> g,
                                      int32_t *__restrict__ dist,
                                      int32_t start, int32_t end,
                                      int32_t *updated) {
    uint32_t upd = 0;
    const int32_t waves = gridDim.x * (blockDim.x / kWave);
    const int32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    for (int32_t u = start + wid; u < end; u += waves) {
        uint32_t gathered = 0;
        int32_t m = group_min<64>(row_min<64, false>(g, unstaged(dist), unstaged<uint32_t>(nullptr),
                                                     g.rbeg(u), g.rend(u), lane_id(), gathered));
        if (lane_id() == 0 && m < dist[u]) {
            dist[u] = m;
            upd++;
        }
    }
    if (lane_id() == 0 && upd) atomicAdd(updated, (int32_t)upd);
}

template <typename OffT>
__global__ void sssp_pull_block_kernel(GraphView<OffT> g,
                                       int32_t *__restrict__ dist,
                                       int32_t start, int32_t end,
                                       int32_t *updated) {
    uint32_t upd = 0;
    for (int32_t u = start + blockIdx.x; u < end; u += gridDim.x) {
        const OffT b = g.rbeg(u), e = g.rend(u);
        int32_t m = HC_INF_DIST;
        for (OffT i = b + threadIdx.x; i < e; i += blockDim.x)
            m = min(m, relax(dist[g.adj[i]], g.wt[i]));
        m = block_min(m);
        if (threadIdx.x == 0 && m < dist[u]) {
            dist[u] = m;
            upd++;
        }
    }
    if (threadIdx.x == 0 && upd) atomicAdd(updated, (int32_t)upd);
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void sssp_pull_fused_kernel(
    PlanDev p, GraphView<OffT> g, int32_t *__restrict__ dist,
    int32_t *updated, int32_t nstage) {
    extern __shared__ int32_t s_stage[];
    uint32_t upd = 0;
    unsigned long long eacc = 0;
    uint32_t gathered = 0;
    sssp_pull_sweep<false, false>(p, g, stage(dist, nstage, s_stage), unstaged<uint32_t>(nullptr),
                                  dist, nullptr, nullptr, upd, eacc, gathered);
    unsigned long long s = block_sum_u64(upd);
    if (threadIdx.x == 0 && s) atomicAdd(updated, (int32_t)s);
}

// Epoch API sweep over the column blocks of a whole graph; `seen` (n bits,
// zeroed by the host) counts each lowered row once.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void sssp_pull_cb_kernel(
    PlanDev plight, PlanDev pcb, GraphView<OffT> g, GraphView<OffT> cbv, int cbk,
    int32_t *__restrict__ dist, int32_t *updated, uint32_t *seen) {
    uint32_t upd = 0, gathered = 0;
    unsigned long long eacc = 0;
    sssp_pull_sweep_cb(pcb, g, cbv, cbk, unstaged(dist), dist, nullptr, seen, upd, eacc,
                       gathered);
    sssp_pull_sweep<false, false>(plight, g, unstaged(dist), unstaged<uint32_t>(nullptr), dist,
                                  nullptr, nullptr, upd, eacc, gathered);
    unsigned long long s = block_sum_u64(upd);
    if (threadIdx.x == 0 && s) atomicAdd(updated, (int32_t)s);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------


// Share of all edges held by the 1% highest-degree rows (cached; the caller
// holds g->mu).  Skewed graphs (Kronecker: about half) and flat ones
// (uniform: about 1%) want different push/pull switch points: a push epoch
// on a flat graph relaxes edges whose targets are spread over the whole
// distance array, so its atomics cost far more per edge.
static double hub_share(hc_graph_s *g) {
    if (g->hub_share >= 0.0) return g->hub_share;
    const int32_t rows = g->rows;
    if (rows <= 0 || g->m <= 0) return g->hub_share = 0.0;
    const int32_t k = std::max<int32_t>(1, rows / 100);
    int64_t top = 0;
    if (g->sorted) {
        top = g->h_off[k] - g->h_off[0];
    } else {
        std::vector<int64_t> deg(rows);
        for (int32_t i = 0; i < rows; i++) deg[i] = g->h_off[i + 1] - g->h_off[i];
        std::nth_element(deg.begin(), deg.begin() + (k - 1), deg.end(), std::greater<int64_t>());
        for (int32_t i = 0; i < k; i++) top += deg[i];
    }
    return g->hub_share = double(top) / double(g->m);
}

bool flat_graph(hc_graph_s *g) { return hub_share(g) < 0.05; }

struct SsspWs {
    FrontierWs f;
    uint16_t *d16 = nullptr;  // 16-bit distance mirror, allocated on first use
};

void free_sssp_ws(hc_graph_s *g) {
    if (!g->ws_sssp) return;
    auto *ws = static_cast<SsspWs *>(g->ws_sssp);
    frontier_ws_free(&ws->f);
    if (ws->d16) hipFree(ws->d16);
    delete ws;
    g->ws_sssp = nullptr;
}

static int get_sssp_ws(hc_graph_s *g, SsspWs **out) {
    // The partial slots are sized for the launch grid; re-allocate when the
    // tuned grid changes.
    if (g->ws_sssp && static_cast<SsspWs *>(g->ws_sssp)->f.grid != grid_blocks()) free_sssp_ws(g);
    if (!g->ws_sssp) {
        auto *ws = new SsspWs();
        int rc = frontier_ws_alloc(g, &ws->f);
        if (rc != HC_OK) {
            delete ws;
            return rc;
        }
        g->ws_sssp = ws;
    }
    *out = static_cast<SsspWs *>(g->ws_sssp);
    return HC_OK;
}

// Bytes a bucketed pull sweep reads and writes per the algorithm: row
// offsets, neighbour id + weight + gathered distance per edge, own distance.
long long sssp_sweep_bytes(const hc_graph_s *g, const PlanHost *ph) {
    const PlanDev &p = ph->dev;
    long long rows = 0;
    for (int c = 1; c < kNumClasses; c++) rows += p.cnt[c];
    rows += p.num_hubs;
    const long long ob = g->off64 ? 8 : 4;
    return ob * (rows + 1) + 12ll * g->m + 4ll * rows;
}

template <typename OffT>
static int launch_pull_epoch(hc_graph_s *g, int mapping, int32_t start,
                             int32_t end, int32_t *dist, int32_t *updated,
                             int blocks, int threads, hipStream_t s) {
    GraphView<OffT> v = view<OffT>(g);
    switch (mapping) {
        case HC_MAP_THREAD:
            if (!threads) threads = 256;
            if (!blocks) blocks = std::max(1, std::min(grid_blocks(), (end - start + threads - 1) / threads));
            sssp_pull_thread_kernel<OffT><<<blocks, threads, 0, s>>>(v, dist, start, end, updated);
            break;
        case HC_MAP_WAVE:
            if (!threads) threads = 256;
            if (!blocks) blocks = grid_blocks();
            sssp_pull_wave_kernel<OffT><<<blocks, threads, 0, s>>>(v, dist, start, end, updated);
            break;
        case HC_MAP_BLOCK:
            if (!threads) threads = 256;
            if (!blocks) blocks = grid_blocks();
            sssp_pull_block_kernel<OffT><<<blocks, threads, 0, s>>>(v, dist, start, end, updated);
            break;
        default: {
            const PlanHost *ph = nullptr;
            // A whole-graph sweep gathers column block by column block.
            const ColBlocks *cb = nullptr;
            if (start == 0 && end == g->n) HC_TRY(get_col_blocks(g, col_shift_for(g, flat_graph(g)), s, &cb));
            if (cb) {
                const PlanHost *pl = nullptr;
                HC_TRY(get_plan(g, 0, g->n, cb_plan_config(cb->K), &ph));
                HC_TRY(get_plan(g, 0, g->n, cb_light_config(cb->K), &pl));
                const int units = ph->total_units * cb->K + pl->total_units;
                if (units == 0) return HC_OK;
                GraphView<OffT> cbv{static_cast<const OffT *>(cb->off), cb->adj, cb->wt, 0, g->n};
                HC_HIP(hipMemsetAsync(cb->seen, 0, sizeof(uint32_t) * ((g->n + 31) / 32), s));
                const int nb = std::max(1, std::min(units, grid_blocks()));
                sssp_pull_cb_kernel<OffT><<<nb, kBlock, 0, s>>>(pl->dev, ph->dev, v, cbv, cb->K,
                                                                dist, updated, cb->seen);
                break;
            }
            HC_TRY(get_plan(g, start, end, sssp_plan_config(), &ph));
            if (ph->total_units == 0) return HC_OK;
            int nb = std::max(1, std::min(ph->total_units, grid_blocks()));
            const int32_t ns = std::min<int32_t>(tuning().sssp_stage, g->n);
            sssp_pull_fused_kernel<OffT><<<nb, kBlock, ns * sizeof(int32_t), s>>>(
                ph->dev, v, dist, updated, ns);
        }
    }
    HC_HIP(hipGetLastError());
    return HC_OK;
}

static int check_launch(int mapping, int blocks, int threads) {
    HC_CHECK_ARG(mapping >= HC_MAP_THREAD && mapping <= HC_MAP_BUCKETED,
                 "unknown mapping %d", mapping);
    HC_CHECK_ARG(blocks >= 0 && threads >= 0, "negative launch shape");
    HC_CHECK_ARG(threads <= 1024, "threads must be <= 1024");
    if (mapping == HC_MAP_WAVE || mapping == HC_MAP_BLOCK)
        HC_CHECK_ARG(threads % kWave == 0,
                     "thread count must be a multiple of 64 for this mapping");
    return HC_OK;
}


template <typename OffT>
struct SsspLaunch {
    PlanDev plan;
    GraphView<OffT> g;
    int32_t *dist;
    uint16_t *d16;  // 16-bit distance mirror (sssp.dist16), or nullptr
    FrontierBufs fb;
    StepCtl *ctl;
    int grid;
    hipStream_t s;
    int32_t nstage;
    int32_t lds_entries;
    int32_t push_cap;
    int defer_dedup;
    KernelTimer *kt;
    PlanDev cbplan;          // column-blocked pull sweeps (cbk > 0)
    GraphView<OffT> cbv;
    int cbk;
    template <bool D16, bool CB>
    void step(FrontierBufs fb, StepCtl *cur) {
        timed_launch(kt, sssp_step_kernel<OffT, D16, CB>, dim3(grid), dim3(kBlock),
                     uint32_t(lds_entries * sizeof(int32_t)), s, plan, g, dist, d16, fb, cur,
                     nstage, push_cap, defer_dedup, cbplan, cbv, cbk);
    }
    // Epoch k: header slot k & 1, counter set k % 3.
    static int launch(void *arg, long long k) {
        auto *a = static_cast<SsspLaunch *>(arg);
        FrontierBufs fb = a->fb;
        fb.part = sum_set(a->fb.part, k);
        StepCtl *cur = a->ctl + (k & 1), *nxt = a->ctl + ((k + 1) & 1);
        if (a->d16) {
            if (a->cbk) a->template step<true, true>(fb, cur);
            else a->template step<true, false>(fb, cur);
        } else {
            if (a->cbk) a->template step<false, true>(fb, cur);
            else a->template step<false, false>(fb, cur);
        }
        sssp_between_kernel<OffT><<<compact_grid(a->grid), kBlock, 0, a->s>>>(
            a->g, fb, cur, nxt, a->ctl, sum_set(a->fb.part, k + 1), a->dist, a->d16);
        if (hipGetLastError() != hipSuccess) return fail(HC_ERR_HIP, "SSSP epoch launch failed");
        return HC_OK;
    }
};


// Bucket width of the single-source loop (knob sssp.delta): 2 x the sampled
// largest weight / mean degree on flat graphs (uniform 23, weights 1..255,
// degree 32: 16; 8 and 16 measured best, 20.7 -> 10.5 ms), off on skewed
// ones (Kronecker 23: 5.1 -> 8.3..9.2 ms with widths 16..256).
static int sssp_delta(hc_graph_s *g) {
    const int d = tuning().sssp_delta;
    if (d >= 0) return d;
    if (!g->weighted || g->m == 0 || g->max_weight <= 0 || !flat_graph(g)) return 0;
    const double mean_deg = (double)g->m / std::max(1, g->rows);
    return (int)std::max(1.0, std::ceil(2.0 * g->max_weight / mean_deg));
}

template <typename OffT>
static int run_sssp_loop(hc_graph_s *g, int32_t src, const int32_t *init_dist,
                         int32_t *dist, int direction, int alpha,
                         int max_epochs, hc_run_stats_t *stats, hipStream_t s) {
    SsspWs *ws = nullptr;
    HC_TRY(get_sssp_ws(g, &ws));
    const PlanHost *ph = nullptr;
    HC_TRY(get_plan(g, 0, g->n, sssp_plan_config(), &ph));
    GraphView<OffT> v = view<OffT>(g);
    if (alpha <= 0) alpha = tuning().sssp_alpha > 0 ? tuning().sssp_alpha : (flat_graph(g) ? 2 : 4);
    uint16_t *d16 = nullptr;
    if (tuning().sssp_dist16 == 1 || (tuning().sssp_dist16 == 2 && flat_graph(g))) {
        if (!ws->d16 && hipMalloc(&ws->d16, sizeof(uint16_t) * std::max(g->n, 1)) != hipSuccess)
            return fail(HC_ERR_OOM, "16-bit distance mirror allocation failed");
        d16 = ws->d16;
    }
    const long long sweep = sssp_sweep_bytes(g, ph);
    const int32_t nstage = std::min<int32_t>(tuning().sssp_stage, g->n);
    const int32_t push_cap = tuning().push_stage;
    const int32_t lds_entries = std::max<int32_t>(nstage, push_cap);
    // Pull epochs gather column block by column block (not with filtering,
    // which tests a changed bit per neighbour before its gather).
    // Buckets (sssp.delta): push epochs only.
    const int delta = init_dist || direction == 1 ? 0 : sssp_delta(g);
    const ColBlocks *cb = nullptr;
    const PlanHost *phb = nullptr;
    if (direction != 2 && delta == 0 && tuning().sssp_filter != 1) {
        HC_TRY(get_col_blocks(g, col_shift_for(g, flat_graph(g)), s, &cb));
        if (cb) {
            // Plain sweep for the light rows only.
            HC_TRY(get_plan(g, 0, g->n, cb_light_config(cb->K), &ph));
            HC_TRY(get_plan(g, 0, g->n, cb_plan_config(cb->K), &phb));
        }
    }
    const void *step_fn =
        d16 ? (cb ? reinterpret_cast<const void *>(&sssp_step_kernel<OffT, true, true>)
                  : reinterpret_cast<const void *>(&sssp_step_kernel<OffT, true, false>))
            : (cb ? reinterpret_cast<const void *>(&sssp_step_kernel<OffT, false, true>)
                  : reinterpret_cast<const void *>(&sssp_step_kernel<OffT, false, false>));
    const int grid = resident_grid(step_fn,
                                   kBlock, lds_entries * sizeof(int32_t),
                                   std::min(ws->f.grid, grid_blocks(tuning().sssp_grid_per_cu)));
    const int init_grid = std::max(1, std::min(grid, (g->n + kBlock - 1) / kBlock));
    sssp_init_kernel<OffT><<<init_grid, kBlock, 0, s>>>(
        v, dist, init_dist, ws->f.fb, ws->f.ctl, src, direction, alpha,
        max_epochs, (long long)g->m, sweep, grid, tuning().sssp_filter, tuning().push_chunk, d16,
        delta);
    HC_HIP(hipGetLastError());
    ws->f.kt.reset();
    // Push epochs deduplicate improved vertices when staging them (flat
    // graphs: few repeats) or on the spot (skewed graphs: hubs' neighbours
    // are lowered by many edges, and repeats would fill the staging).
    const int defer_dedup = tuning().sssp_push_dedup == 2 ? (flat_graph(g) ? 1 : 0)
                                                          : tuning().sssp_push_dedup;
    SsspLaunch<OffT> L{ph->dev, v, dist, d16, ws->f.fb, ws->f.ctl, grid, s, nstage, lds_entries,
                       push_cap, defer_dedup, &ws->f.kt, phb ? phb->dev : PlanDev{},
                       cb ? GraphView<OffT>{static_cast<const OffT *>(cb->off), cb->adj, cb->wt,
                                            0, g->n}
                          : GraphView<OffT>{},
                       cb ? cb->K : 0};
    // Bellman-Ford needs at most n epochs (buckets add promotion epochs and
    // re-pushes: a generous bound); the cap bounds a faulty loop.
    const long long bound = delta > 0 ? 64LL * (g->n + 1) + 4096 : (long long)g->n;
    const long long cap = (max_epochs > 0 ? std::min<long long>(max_epochs, bound) : bound) + 2;
    float ms = 0;
    long long launches = 0;
    StepCtl fc;
    HC_TRY(run_epoch_loop(&ws->f, s, cap, &SsspLaunch<OffT>::launch, &L, &ms,
                          &launches, &fc));
    if (!fc.done && max_epochs <= 0)
        return fail(HC_ERR_HIP, "SSSP loop did not converge in %lld epochs", launches);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        // Device wall clock from the init kernel to the controller of the
        // last epoch (launches queued past it are not part of the loop).
        stats->ms = double(fc.t_end - fc.t_start) / device_state().clock_khz;
        stats->epochs = fc.epoch;
        stats->push_epochs = fc.push_epochs;
        stats->pull_epochs = fc.pull_epochs;
        stats->launches = (int32_t)(2 * launches);
        stats->reached = -1;
        stats->alg_bytes = (double)fc.alg_bytes;
        if (ws->f.kt.on) {
            stats->kernel_launches = std::min(fc.epoch, ws->f.kt.pairs());
            stats->kernel_ms = ws->f.kt.sum_ms(stats->kernel_launches);
        }
    }
    return HC_OK;
}

}  // namespace hc

using namespace hc;

extern "C" {

int hc_sssp_pull_epoch(hc_graph_t g, int mapping, int32_t start_id,
                       int32_t end_id, int32_t *dist, int32_t *updated,
                       int blocks, int threads, hc_stream_t stream) {
    HC_CHECK_ARG(g && dist && updated, "NULL argument");
    HC_CHECK_ARG(g->weighted, "SSSP needs a weighted graph");
    HC_CHECK_ARG(g->row0 <= start_id && start_id <= end_id &&
                     end_id <= g->row0 + g->rows,
                 "range [%d,%d) outside the rows held [%d,%d)", start_id,
                 end_id, g->row0, g->row0 + g->rows);
    HC_TRY(check_launch(mapping, blocks, threads));
    HC_TRY(ensure_init());
    if (start_id == end_id) return HC_OK;
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    return g->off64 ? launch_pull_epoch<uint64_t>(g, mapping, start_id, end_id, dist, updated, blocks, threads, s)
                    : launch_pull_epoch<uint32_t>(g, mapping, start_id, end_id, dist, updated, blocks, threads, s);
}

int hc_sssp(hc_graph_t g, int32_t source, int32_t *dist,
            const hc_sssp_opts_t *opts, hc_run_stats_t *stats,
            hc_stream_t stream) {
    HC_CHECK_ARG(g && dist, "NULL argument");
    HC_CHECK_ARG(g->weighted, "SSSP needs a weighted graph");
    HC_CHECK_ARG(g->row0 == 0 && g->rows == g->n,
                 "full runs need a whole-graph handle");
    HC_CHECK_ARG(source >= 0 && source < g->n, "source %d outside [0,%d)",
                 source, g->n);
    int direction = opts ? opts->direction : 0;
    int alpha = opts && opts->alpha > 0 ? opts->alpha : 0;  // 0: the loop's default
    int max_epochs = opts ? opts->max_epochs : 0;
    HC_CHECK_ARG(direction >= 0 && direction <= 2, "direction must be 0, 1 or 2");
    HC_TRY(ensure_init());
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    return g->off64 ? run_sssp_loop<uint64_t>(g, source, nullptr, dist, direction, alpha, max_epochs, stats, s)
                    : run_sssp_loop<uint32_t>(g, source, nullptr, dist, direction, alpha, max_epochs, stats, s);
}

int hc_sssp_pull(hc_graph_t g, int mapping, const int32_t *init_dist,
                 int32_t *dist, int blocks, int threads,
                 hc_run_stats_t *stats, hc_stream_t stream) {
    HC_CHECK_ARG(g && init_dist && dist, "NULL argument");
    HC_CHECK_ARG(g->weighted, "SSSP needs a weighted graph");
    HC_CHECK_ARG(g->row0 == 0 && g->rows == g->n,
                 "full runs need a whole-graph handle");
    HC_TRY(check_launch(mapping, blocks, threads));
    HC_TRY(ensure_init());
    hipStream_t s = pick_stream(stream);
    if (mapping == HC_MAP_BUCKETED) {
        GraphCall call(g);
        return g->off64 ? run_sssp_loop<uint64_t>(g, 0, init_dist, dist, 1, 8, 0, stats, s)
                        : run_sssp_loop<uint32_t>(g, 0, init_dist, dist, 1, 8, 0, stats, s);
    }
    // Host-driven loop of sssp_pull_gpu (src/kernels/gpu/sssp.cuh:76-85):
    // zero the counter, run one epoch, read the counter back.
    int32_t *d_upd = nullptr;
    HC_HIP(hipMalloc(&d_upd, sizeof(int32_t)));
    int rc = HC_OK;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    if (hipMemcpyAsync(dist, init_dist, sizeof(int32_t) * g->n,
                       hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = fail(HC_ERR_HIP, "copy of init_dist failed");
    int epochs = 0;
    hipEventRecord(e0, s);
    int32_t h_upd = 1;
    while (rc == HC_OK && h_upd != 0) {
        hipMemsetAsync(d_upd, 0, sizeof(int32_t), s);
        {
            GraphCall call(g);
            rc = g->off64 ? launch_pull_epoch<uint64_t>(g, mapping, 0, g->n, dist, d_upd, blocks, threads, s)
                          : launch_pull_epoch<uint32_t>(g, mapping, 0, g->n, dist, d_upd, blocks, threads, s);
        }
        if (rc != HC_OK) break;
        if (hipMemcpyAsync(&h_upd, d_upd, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = fail(HC_ERR_HIP, "epoch failed");
        epochs++;
    }
    hipEventRecord(e1, s);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(d_upd);
    if (rc == HC_OK && stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = ms;
        stats->epochs = epochs;
        stats->pull_epochs = epochs;
        stats->launches = epochs;
    }
    return rc;
}

}  // extern "C"
