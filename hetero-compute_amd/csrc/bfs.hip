// BFS epoch kernels for gfx950.
//
//  * Bottom-up (pull) epoch, semantics of epoch_bfs_pull_cpu
//    (src/kernels/cpu/bfs.cuh:305-335): every unvisited vertex takes the
//    first neighbour (CSR order) found in the frontier bitmap as its parent.
//    Lane groups of 1/4/16/64 test consecutive neighbours and stop at the
//    first ballot hit, so the parent chosen is the same as the sequential
//    scan's.
//  * Top-down (push) epoch, semantics of epoch_bfs_push_cpu_by_node
//    (src/kernels/cpu/bfs.cuh:212-244): frontier vertices claim unvisited
//    neighbours with compare-and-swap; winners are appended to the next
//    frontier (wavefront-aggregated) and their degrees summed (scout count).
//  * Direction-optimizing loop, switching rule of bfs_do_cpu<alpha,beta>
//    (src/kernels/cpu/bfs.cuh:48-114), evaluated on the device by the
//    between-epochs launch.  Bottom-up epochs emit the next frontier as a
//    bitmap only; top-down epochs emit bitmap and queues; the same launch
//    builds queues from the bitmap when the direction turns from bottom-up
//    to top-down.
//  * A bottom-up epoch lists the vertices it leaves unvisited (per-workgroup
//    list segments, no global counter); the next bottom-up epoch sweeps only
//    those.  One-lane rows read their first neighbour from a per-vertex
//    array, so the common first-probe hit costs coalesced loads only.
#include <algorithm>
#include <cstring>
#include <vector>

#include "frontier.h"

namespace hc {

// Frontier bitmap words, the first ones (the hubs' bits) from an LDS copy;
// the bitmap is read-only during a bottom-up epoch, so the copy is exact.
__device__ __forceinline__ bool bit_test(const Staged<uint32_t> &bm, int32_t v) {
    return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Scans row [b, e) of u with G lanes from lane `lane` (group base lane gbase
// within the wavefront); returns the first neighbour in the frontier (CSR
// order).  The first round reads one position per lane: neighbours are in
// id order, ids in degree order, and the low-id hubs are the likeliest
// parents, so most scans end there.  Later rounds read four positions per
// lane (predicated) and test their bits together; the first hit is the
// earliest position of the earliest round with one.  `scanned` counts the
// positions examined.  FIRST (one lane): the first neighbour is `first_v`,
// already read from the first-neighbour array (-1: empty row).
template <int G, bool FIRST = false, typename OffT>
__device__ __forceinline__ bool bu_scan(const GraphView<OffT> &g,
                                        const Staged<uint32_t> &frontier,
                                        OffT b, OffT e, int lane, int gbase,
                                        int32_t &parent, uint32_t &scanned,
                                        int32_t first_v = -1) {
    static_assert(!FIRST || G == 1, "first-neighbour probes are one-lane");
    const int L = G == 1 ? 0 : lane;
    {
        const OffT k = b + OffT(L);
        const int32_t v = FIRST ? first_v : (k < e ? ldnt(&g.adj[k]) : -1);
        const bool h = v >= 0 && bit_test(frontier, v);
        if (G == 1) {
            if (v < 0) return false;
            scanned++;
            if (h) {
                parent = v;
                return true;
            }
        } else {
            scanned += v >= 0 ? 1u : 0u;
            const uint64_t m = __ballot(h);
            const uint64_t gm = G == 64 ? m : ((m >> gbase) & ((uint64_t(1) << G) - 1));
            if (gm) {
                parent = __shfl(v, gbase + __ffsll((unsigned long long)gm) - 1);
                return true;
            }
        }
    }
    for (OffT base = b + OffT(G); base < e; base += OffT(4 * G)) {
        int32_t v[4];
        bool h[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const OffT k = base + OffT(r * G + L);
            v[r] = k < e ? ldnt(&g.adj[k]) : -1;
        }
#pragma unroll
        for (int r = 0; r < 4; r++) h[r] = v[r] >= 0 && bit_test(frontier, v[r]);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (G == 1) {
                if (v[r] < 0) return false;
                scanned++;
                if (h[r]) {
                    parent = v[r];
                    return true;
                }
            } else {
                scanned += v[r] >= 0 ? 1u : 0u;
                const uint64_t m = __ballot(h[r]);
                const uint64_t gm = G == 64 ? m : ((m >> gbase) & ((uint64_t(1) << G) - 1));
                if (gm) {
                    parent = __shfl(v[r], gbase + __ffsll((unsigned long long)gm) - 1);
                    return true;
                }
            }
        }
    }
    return false;
}

// Bottom-up step of one vertex uu by a group of G lanes.  found: a frontier
// neighbour was found (parent par); open: uu is unvisited and stays so (it
// has no frontier neighbour).  Results are valid in the group's lane 0.
// With `first` (one lane): the visited flag, the first neighbour and the row
// bounds are read together, so most rows cost one round of coalesced loads
// and a bitmap probe instead of a dependent gather into the edge array.
// deg: the row's degree when its bounds were read here (-1 otherwise), so the
// commit does not read them again.
template <int G, bool CHECK, typename OffT>
__device__ __forceinline__ void bu_vertex(const GraphView<OffT> &g,
                                          const int32_t *__restrict__ visited,
                                          const int32_t *__restrict__ first,
                                          const Staged<uint32_t> &frontier, int32_t uu,
                                          int lane, int gbase, bool &found, bool &open,
                                          int32_t &par, uint32_t &scanned, int64_t &deg) {
    if (G == 1 && first) {
        const int32_t vis = CHECK ? visited[uu] : -1;
        const int32_t f = first[uu];
        const OffT b = g.rbeg(uu), e = g.rend(uu);
        deg = int64_t(e - b);
        if (vis < 0) {
            int32_t pr = -1;
            const bool fnd = bu_scan<1, true>(g, frontier, b, e, 0, 0, pr, scanned, f);
            found = fnd;
            open = !fnd;
            par = pr;
        }
        return;
    }
    if (G > 1 && first) {
        // The group tests the first neighbour from the array (one broadcast
        // load) and scans the row from its second edge only if that misses.
        const int32_t vis = CHECK ? visited[uu] : -1;
        const int32_t f = first[uu];
        if (vis < 0) {
            int32_t pr = f;
            bool fnd = f >= 0 && bit_test(frontier, f);
            if (!fnd) {
                pr = -1;
                fnd = bu_scan<G>(g, frontier, g.rbeg(uu) + OffT(1), g.rend(uu), lane, gbase, pr,
                                 scanned);
            }
            if (lane == 0) {
                scanned++;
                found = fnd;
                open = !fnd;
                par = pr;
            }
        }
        return;
    }
    if (!CHECK || visited[uu] < 0) {
        int32_t pr = -1;
        const OffT b = g.rbeg(uu), e = g.rend(uu);
        deg = int64_t(e - b);
        bool f = bu_scan<G>(g, frontier, b, e, lane, gbase, pr, scanned);
        if (lane == 0) {
            found = f;
            open = !f;
            par = pr;
        }
    }
}

// A one-lane row of the plan's last class read ahead of its use (visited
// flag, first neighbour, bounds), and its bottom-up step once resolved:
// the same probe and scan as bu_vertex<1> with first-neighbour probes.
template <typename OffT>
struct RowAhead {
    int32_t u, vis, f;
    OffT b, e;
    bool ok;
};

template <typename OffT>
__device__ __forceinline__ RowAhead<OffT> row_ahead(const PlanDev &p, int k,
                                                    const GraphView<OffT> &g,
                                                    const int32_t *__restrict__ visited,
                                                    const int32_t *__restrict__ first) {
    RowAhead<OffT> r;
    const int idx = k * kBlock + threadIdx.x;
    r.ok = idx < p.cnt[5];
    r.u = r.ok ? (p.list[5] ? p.list[5][idx] : p.beg[5] + idx) : 0;
    r.vis = r.ok ? visited[r.u] : 0;
    r.f = r.ok ? first[r.u] : -1;
    r.b = r.ok ? g.rbeg(r.u) : OffT(0);
    r.e = r.ok ? g.rend(r.u) : OffT(0);
    return r;
}

template <typename OffT>
__device__ __forceinline__ void row_resolve(const RowAhead<OffT> &r, const GraphView<OffT> &g,
                                            const Staged<uint32_t> &frontier, bool &found,
                                            bool &open, int32_t &par, uint32_t &scanned,
                                            int64_t &deg) {
    deg = int64_t(r.e - r.b);
    if (!r.ok || r.vis >= 0) return;
    int32_t pr = -1;
    const bool fnd = bu_scan<1, true>(g, frontier, r.b, r.e, 0, 0, pr, scanned, r.f);
    found = fnd;
    open = !fnd;
    par = pr;
}

template <int G, typename OffT>
__device__ __forceinline__ void bu_group(const PlanDev &p, int c, int k,
                                         const GraphView<OffT> &g,
                                         const int32_t *__restrict__ visited,
                                         const int32_t *__restrict__ first,
                                         const Staged<uint32_t> &frontier,
                                         bool &found, bool &open, int32_t &u, int32_t &par,
                                         uint32_t &scanned, int64_t &deg) {
    constexpr int PER = kBlock / G;
    const int gid = threadIdx.x / G, lane = threadIdx.x % G;
    const int gbase = lane_id() - lane;
    const int idx = k * PER + gid;
    if (idx < p.cnt[c]) {
        int32_t uu = p.list[c] ? p.list[c][idx] : p.beg[c] + idx;
        u = uu;
        bu_vertex<G, true>(g, visited, first, frontier, uu, lane, gbase, found, open, par,
                           scanned, deg);
    }
}

// Records the outcome of one bottom-up vertex: LOOP: depth/parent, a bit in
// the next-frontier bitmap (counted with its degree) and, with `rem`, the
// vertex on the next unvisited list when it stays open.  Epoch API: visited
// is parents.  Called by every lane of the wavefront (wave-aggregated).
template <bool LOOP, typename OffT>
__device__ __forceinline__ void bu_commit(const GraphView<OffT> &g,
                                          int32_t *__restrict__ visited,
                                          int32_t *__restrict__ parents, int32_t level,
                                          uint32_t *__restrict__ next, int32_t *rem,
                                          uint32_t *rem_cnt, bool found, bool open,
                                          int32_t u, int32_t par, int64_t deg, uint32_t &cnt,
                                          unsigned long long &edges_acc) {
    if (LOOP) {
        if (found) {
            visited[u] = level;
            if (parents) parents[u] = par;
        }
        // Only u's own lane sets u's bit in a bottom-up epoch.
        wave_or_bits(next, u, found);
        if (found) {
            cnt++;
            edges_acc += (unsigned long long)(deg >= 0 ? deg : int64_t(g.rend(u) - g.rbeg(u)));
        }
        if (rem) wave_append(rem, rem_cnt, open, u);
    } else if (found) {
        visited[u] = par;
        atomicOr(&next[u >> 5], 1u << (u & 31));
        cnt++;
    }
}

// One bottom-up sweep over the plan (classes 2..5; the BFS plan keeps classes
// 0 and 1 empty).  LOOP: visited is depth; found vertices get depth/parent
// and a bit in the next-frontier bitmap, counted with their degree; with
// `rem`, vertices that stay unvisited are listed for the next sweep.
// Epoch API: visited is parents.
template <bool LOOP, typename OffT>
__device__ void bu_sweep(const PlanDev &p, const GraphView<OffT> &g,
                         int32_t *__restrict__ visited, int32_t *__restrict__ parents,
                         const int32_t *__restrict__ first, const int32_t *__restrict__ first_g,
                         int32_t level, const Staged<uint32_t> &frontier,
                         uint32_t *__restrict__ next, int32_t *rem, uint32_t *rem_cnt,
                         uint32_t &cnt, unsigned long long &edges_acc, uint32_t &scanned) {
    const int total = p.unit_start[kNumClasses];
    for (int unit = blockIdx.x; unit < total; unit += gridDim.x) {
        int c = 0;
        while (unit >= p.unit_start[c + 1]) c++;
        const int k = unit - p.unit_start[c];
        if (LOOP && c == 5 && first) {
            // One-lane rows, the last class: this workgroup's remaining
            // units in a pipeline that reads the next unit's rows before
            // committing this unit's, so the commit's stores and atomics
            // are not waited for before the next loads are in flight.
            RowAhead<OffT> cur = row_ahead(p, k, g, visited, first);
            for (;;) {
                const int nunit = unit + (int)gridDim.x;
                const bool more = nunit < total;
                RowAhead<OffT> nx{0, 0, -1, OffT(0), OffT(0), false};
                if (more) nx = row_ahead(p, nunit - p.unit_start[5], g, visited, first);
                bool found = false, open = false;
                int32_t par = -1;
                int64_t deg = -1;
                row_resolve(cur, g, frontier, found, open, par, scanned, deg);
                bu_commit<LOOP>(g, visited, parents, level, next, rem, rem_cnt, found, open,
                                cur.u, par, deg, cnt, edges_acc);
                if (!more) return;
                unit = nunit;
                cur = nx;
            }
        }
        bool found = false, open = false;
        int32_t u = 0, par = -1;
        int64_t deg = -1;
        if (c == 2)
            bu_group<64>(p, c, k, g, visited, first_g, frontier, found, open, u, par, scanned,
                         deg);
        else if (c == 3)
            bu_group<16>(p, c, k, g, visited, first_g, frontier, found, open, u, par, scanned,
                         deg);
        else if (c == 4)
            bu_group<4>(p, c, k, g, visited, first_g, frontier, found, open, u, par, scanned,
                        deg);
        else if (c == 5)
            bu_group<1>(p, c, k, g, visited, first, frontier, found, open, u, par, scanned, deg);
        bu_commit<LOOP>(g, visited, parents, level, next, rem, rem_cnt, found, open, u, par, deg,
                        cnt, edges_acc);
    }
}

// Unvisited lists of the bottom-up loop.  A pull epoch lists, per
// workgroup, the plan vertices it left unvisited; workgroup b owns the slots
// [off[b], off[b+1]) of each parity's list (as many as plan vertices it
// sweeps, so the lists need no global counter) and records its list length
// in len[parity * grid + b].  A following pull epoch sweeps only the lists.
struct RemLists {
    const int32_t *off;
    uint32_t *len;
};

// Bottom-up sweep of this workgroup's unvisited list, G lanes per vertex;
// vertices that stay open go on the next list.
template <int G, typename OffT>
__device__ void bu_list_sweep(const GraphView<OffT> &g, int32_t *__restrict__ depth,
                              int32_t *__restrict__ parents,
                              const int32_t *__restrict__ first, int32_t level,
                              const Staged<uint32_t> &frontier, uint32_t *__restrict__ next,
                              const int32_t *__restrict__ list, uint32_t len, int32_t *rem,
                              uint32_t *rem_cnt, uint32_t &cnt, unsigned long long &edges_acc,
                              uint32_t &scanned) {
    constexpr int PER = kBlock / G;
    const int gid = threadIdx.x / G, lane = threadIdx.x % G;
    const int gbase = lane_id() - lane;
    // Listed vertices are unvisited: only pull epochs ran since they were
    // listed, and each visits only the vertices it scans.
    for (uint32_t base = 0; base < len; base += PER) {
        const uint32_t idx = base + gid;
        bool found = false, open = false;
        int32_t u = 0, par = -1;
        int64_t deg = -1;
        if (idx < len) {
            u = list[idx];
            bu_vertex<G, false>(g, depth, G == 1 ? first : nullptr, frontier, u, lane, gbase,
                                found, open, par, scanned, deg);
        }
        bu_commit<true>(g, depth, parents, level, next, rem, rem_cnt, found, open, u, par, deg,
                        cnt, edges_acc);
    }
}

// Top-down expansion of one frontier row slice.
template <int G, typename OffT>
__device__ __forceinline__ void td_range(const GraphView<OffT> &g,
                                         int32_t *__restrict__ depth,
                                         int32_t *__restrict__ parents,
                                         int32_t level, uint32_t *__restrict__ next,
                                         const FrontierBufs &fb, StepCtl *ctl,
                                         const PushOut &out,
                                         int nxt, int32_t u, OffT b, OffT e,
                                         int lane, uint32_t &cnt,
                                         unsigned long long &edges_acc,
                                         uint32_t &scanned) {
    for (OffT i = b + OffT(lane); i < e; i += OffT(G)) {
        int32_t v = g.adj[i];
        scanned++;
        bool app = false;
        // The neighbour's bounds are read with its depth (one round trip
        // before the claim instead of one after it).
        const int32_t dv = depth[v];
        const OffT vb = g.rbeg(v), ve = g.rend(v);
        if (dv < 0 && atomicCAS(&depth[v], -1, level) == -1) {
            app = true;
            if (parents) parents[v] = u;
            atomicOr(&next[v >> 5], 1u << (v & 31));
            cnt++;
            edges_acc += (unsigned long long)(ve - vb);
        }
        push_out_append(out, g, fb, ctl, nxt, app, v, cnt, edges_acc);
    }
}

template <typename OffT>
__device__ void td_sweep(const GraphView<OffT> &g, int32_t *__restrict__ depth,
                         int32_t *__restrict__ parents, int32_t level,
                         uint32_t *__restrict__ next, const FrontierBufs &fb,
                         StepCtl *ctl, const PushOut &out, int cur, int nxt, uint32_t &cnt,
                         unsigned long long &edges_acc, uint32_t &scanned) {
    const uint32_t nch = fb.cnt[cur][3], n0 = fb.cnt[cur][0],
                   n1 = fb.cnt[cur][1], n2 = fb.cnt[cur][2];
    const uint32_t u1 = nch + n0;
    const uint32_t u2 = u1 + (n1 + kWavesPerBlock - 1) / kWavesPerBlock;
    const uint32_t total = u2 + (n2 + kSmallPerUnit - 1) / kSmallPerUnit;
    // Edges per unit: a hub chunk, a workgroup row (<= kHubDeg), wavefront
    // rows, or 8-lane rows.
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
            td_range<kBlock>(g, depth, parents, level, next, fb, ctl, out, nxt, h.u,
                             OffT(h.b), OffT(h.e), threadIdx.x, cnt, edges_acc, scanned);
        } else if (unit < u1) {
            int32_t u = fb.q[cur][0][unit - nch];
            td_range<kBlock>(g, depth, parents, level, next, fb, ctl, out, nxt, u,
                             g.rbeg(u), g.rend(u), threadIdx.x, cnt, edges_acc, scanned);
        } else if (unit < u2) {
            uint32_t idx = (unit - u1) * kWavesPerBlock + threadIdx.x / kWave;
            if (idx < n1) {
                int32_t u = fb.q[cur][1][idx];
                td_range<kWave>(g, depth, parents, level, next, fb, ctl, out, nxt, u,
                                g.rbeg(u), g.rend(u), lane_id(), cnt, edges_acc, scanned);
            }
        } else {
            uint32_t idx = (unit - u2) * kSmallPerUnit + threadIdx.x / kSmallLanes;
            if (idx < n2) {
                int32_t u = fb.q[cur][2][idx];
                td_range<kSmallLanes>(g, depth, parents, level, next, fb, ctl, out, nxt,
                                      u, g.rbeg(u), g.rend(u),
                                      threadIdx.x % kSmallLanes, cnt, edges_acc, scanned);
            }
        }
    }
}

// `nstage` leading frontier-bitmap words are staged in dynamic LDS for
// bottom-up epochs; top-down epochs stage up to `push_cap` output entries
// there.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void bfs_step_kernel(
    PlanDev p, GraphView<OffT> g, int32_t *__restrict__ depth,
    int32_t *__restrict__ parents, const int32_t *__restrict__ first, FrontierBufs fb,
    StepCtl *ctl, RemLists rl, int32_t nstage, int32_t stage_min, int32_t push_cap,
    int first_all) {
    extern __shared__ uint32_t s_words[];
    __shared__ uint32_t s_rem;
    if (ctl->done) return;
    const int mode = ctl->mode;
    const int k = ctl->epoch, cur = k & 1, nxt = cur ^ 1;
    const uint32_t *frontier = fb.bm[k % 3];
    uint32_t *next = fb.bm[(k + 1) % 3];
    uint32_t *stale = fb.bm[(k + 2) % 3];
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < fb.words;
         w += gridDim.x * kBlock)
        stale[w] = 0;
    const int32_t level = k + 1;
    uint32_t cnt = 0, scanned = 0;
    unsigned long long eacc = 0, list_in = 0, list_out = 0;
    if (mode == kPull) {
        // The unvisited lists live in the small-degree push queues, which
        // pull epochs do not use (a switch to push invalidates them).
        const bool listing = ctl->rem_knob != 0;
        if (threadIdx.x == 0) s_rem = 0;
        const int32_t off = listing ? rl.off[blockIdx.x] : 0;
        const uint32_t len = ctl->rem_valid ? rl.len[cur * gridDim.x + blockIdx.x] : 0u;
        // A short list does not repay copying the bitmap head into LDS.
        const int32_t ns = (ctl->rem_valid && len < (uint32_t)stage_min) ? 0 : nstage;
        const Staged<uint32_t> fr = stage(frontier, ns, s_words);  // syncs s_rem
        int32_t *rem = listing ? fb.q[nxt][2] + off : nullptr;
        if (ctl->rem_valid) {
            const int32_t *list = fb.q[cur][2] + off;
            if (ctl->rem_knob == 16)
                bu_list_sweep<16>(g, depth, parents, first, level, fr, next, list, len, rem, &s_rem,
                                  cnt, eacc, scanned);
            else if (ctl->rem_knob == 4)
                bu_list_sweep<4>(g, depth, parents, first, level, fr, next, list, len, rem, &s_rem,
                                 cnt, eacc, scanned);
            else
                bu_list_sweep<1>(g, depth, parents, first, level, fr, next, list, len, rem, &s_rem,
                                 cnt, eacc, scanned);
        } else
            bu_sweep<true>(p, g, depth, parents, first, first_all ? first : nullptr, level, fr,
                           next, rem, &s_rem, cnt, eacc, scanned);
        if (listing) {
            __syncthreads();
            if (threadIdx.x == 0) rl.len[nxt * gridDim.x + blockIdx.x] = s_rem;
            list_in = len;
            list_out = s_rem;
        }
    } else {
        // Top-down: output staged in the dynamic LDS (unused by this mode).
        __shared__ uint32_t s_out;
        if (threadIdx.x == 0) s_out = 0;
        __syncthreads();
        // Claims by compare-and-swap are unique: no deduplication.
        const PushOut out{reinterpret_cast<int32_t *>(s_words), &s_out, (uint32_t)push_cap,
                          nullptr};
        td_sweep(g, depth, parents, level, next, fb, ctl, out, cur, nxt, cnt, eacc, scanned);
        push_out_flush(out, g, fb, ctl, nxt);
    }
    write_partials(fb.part, cnt, eacc, scanned, list_in, list_out);
}

// Direction-optimizing controller (one thread, on the header copy c): the
// decisions of bfs_do_cpu (src/kernels/cpu/bfs.cuh:68-103) applied between
// epochs, from the step sums of the epoch that ran.
__device__ void bfs_decide(StepCtl *c, StepCtl *trace, const unsigned long long (&sums)[kSumSlots],
                           int offb) {
    // Frontier size and degree sum, edges examined, unvisited-list lengths
    // read and written by a pull epoch.
    const unsigned long long q = sums[0], sc = sums[1], scanned = sums[2];
    const unsigned long long lin = sums[3], lout = sums[4];
    const int k = c->epoch;
    const int ran = c->mode;
    c->reached += q;
    trace_epoch(c, trace, k, ran, q, sc, scanned);
    if (ran == kPush) {
        // queue entry + offsets per frontier vertex; id + depth probe per
        // edge; depth, parent, bitmap and queue writes per new vertex.
        c->alg_bytes += (unsigned long long)(4 + 2 * offb) * c->in_n + 8ull * scanned + 16ull * q;
        c->push_epochs++;
    } else {
        // visited flag per swept row, id + bitmap bit per examined edge,
        // writes per new vertex, bitmap clear.
        // List mode: a list entry per swept row; list writes.
        const unsigned long long rows = c->rem_valid ? lin : (unsigned long long)c->plan_rows;
        c->alg_bytes += 4ull * rows + 4ull * scanned + scanned / 8 + 12ull * q +
                        4ull * ((c->n + 31) / 32) + 4ull * lout;
        c->pull_epochs++;
    }
    c->epoch = k + 1;
    c->in_n = (long long)q;
    c->in_e = (long long)sc;
    const long long n = c->n;
    int next = ran;
    if (c->direction != 0) {
        next = c->direction == 1 ? kPull : kPush;
        if (q == 0) c->done = 1;
    } else if (ran == kPush) {
        c->scout = (long long)sc;
        if (q == 0) {
            c->done = 1;
        } else if (c->scout > c->edges_to_check / c->alpha) {
            next = kPull;
            c->num_nodes = (long long)q;
        } else {
            next = kPush;
            c->edges_to_check -= c->scout;
            c->scout = 0;
        }
    } else {
        c->prev_nodes = c->num_nodes;
        c->num_nodes = (long long)q;
        if ((long long)q > c->prev_nodes || (long long)q > n / c->beta) {
            next = kPull;
        } else {
            c->scout = 1;
            if (q == 0) {
                c->done = 1;
            } else if (c->scout > c->edges_to_check / c->alpha) {
                next = kPull;
                c->num_nodes = (long long)q;
            } else {
                next = kPush;
                c->edges_to_check -= c->scout;
                c->scout = 0;
            }
        }
    }
    if (c->max_epochs > 0 && k + 1 >= c->max_epochs) c->done = 1;
    c->need_compact = (!c->done && ran == kPull && next == kPush) ? 1 : 0;
    // A pull epoch listed the vertices it left open; the next pull reads that
    // list unless the list mode is switched off.
    c->rem_valid = (c->rem_knob && ran == kPull && next == kPull) ? 1 : 0;
    c->mode = next;
}

// Between epochs k and k + 1: the controller (every workgroup decides on its
// own copy; workgroup 0 records the epoch and stores header k + 1), then
// the frontier compaction when the loop turns to top-down.  A controller
// run by the step's last workgroup was measured slower: every workgroup
// then needs a device-scope release fence, which writes back its XCD's L2.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void bfs_between_kernel(GraphView<OffT> g, FrontierBufs fb,
                                                             StepCtl *cur, StepCtl *nxt,
                                                             StepCtl *trace, Partials next_set,
                                                             int offb) {
    __shared__ StepCtl s_c;  // header fields only
    unsigned long long sums[kSumSlots];
    if (!ctl_begin(&s_c, cur, fb.part, sums)) return;
    if (threadIdx.x == 0) bfs_decide(&s_c, blockIdx.x == 0 ? trace : nullptr, sums, offb);
    __syncthreads();
    ctl_end(&s_c, cur, nxt, fb, next_set);
    compact_body(g, fb, &s_c, 1);
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void bfs_init_kernel(
    GraphView<OffT> g, int32_t *__restrict__ depth, int32_t *__restrict__ parents,
    FrontierBufs fb, StepCtl *ctl, int32_t src, int direction, int alpha,
    int beta, int max_epochs, long long m, int32_t plan_rows, int grid, int rem_knob,
    int chunk) {
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < g.n; i += stride) {
        depth[i] = i == src ? 0 : -1;
        if (parents) parents[i] = i == src ? src : -1;
    }
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < fb.words; w += stride) {
        fb.bm[0][w] = (w == (src >> 5)) ? (1u << (src & 31)) : 0u;
        fb.bm[1][w] = 0;
        fb.bm[2][w] = 0;
    }
    if (blockIdx.x == 0)
        for (int j = 0; j < kSumSets; j++) clear_partials(sum_set(fb.part, j));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int p = 0; p < 2; p++)
            for (int j = 0; j < 5; j++) fb.cnt[p][j] = 0;
        ctl->done = 0;
        ctl->epoch = 0;
        ctl->max_epochs = max_epochs;
        ctl->direction = direction;
        ctl->alpha = alpha;
        ctl->beta = beta;
        ctl->need_compact = 0;
        ctl->rem_valid = 0;
        ctl->rem_knob = rem_knob;
        ctl->chunk = chunk;
        ctl->n = g.n;
        ctl->grid = grid;
        ctl->m = m;
        ctl->push_epochs = ctl->pull_epochs = 0;
        ctl->reached = 1;
        ctl->alg_bytes = 0;
        ctl->t_start = wall_clock64();
        ctl->t_end = ctl->t_start;
        ctl->plan_rows = plan_rows;
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
        // Initial decision of bfs_do_cpu: num_edges = deg(src),
        // edges_to_check = m.
        ctl->edges_to_check = m;
        ctl->scout = d;
        ctl->num_nodes = 1;
        ctl->prev_nodes = 0;
        if (direction == 1) {
            ctl->mode = kPull;
        } else if (direction == 2) {
            ctl->mode = kPush;
        } else if (ctl->scout > ctl->edges_to_check / alpha) {
            ctl->mode = kPull;
        } else {
            ctl->mode = kPush;
            ctl->edges_to_check -= ctl->scout;
            ctl->scout = 0;
        }
    }
    if (blockIdx.x == 0) write_source_chunks(g, fb, src, chunk);
}

// ---- epoch-API kernels (reference semantics: parents marks visited) ----

template <typename OffT>
__global__ __launch_bounds__(kBlock) void bfs_pull_fused_kernel(
    PlanDev p, GraphView<OffT> g, int32_t *__restrict__ parents,
    const uint32_t *__restrict__ frontier, uint32_t *__restrict__ next,
    int32_t *num_nodes, int32_t nstage) {
    extern __shared__ uint32_t s_words[];
    uint32_t nnew = 0, scanned = 0;
    unsigned long long eacc = 0;
    bu_sweep<false>(p, g, parents, nullptr, nullptr, nullptr, 0, stage(frontier, nstage, s_words), next,
                    nullptr, nullptr, nnew, eacc, scanned);
    unsigned long long s = block_sum_u64(nnew);
    if (threadIdx.x == 0 && s) atomicAdd(num_nodes, (int32_t)s);
}

template <typename OffT>
__global__ void bfs_pull_thread_kernel(GraphView<OffT> g,
                                       int32_t *__restrict__ parents,
                                       int32_t start, int32_t end,
                                       const uint32_t *__restrict__ frontier,
                                       uint32_t *__restrict__ next,
                                       int32_t *num_nodes) {
    uint32_t nnew = 0, scanned = 0;
    const int32_t stride = gridDim.x * blockDim.x;
    for (int32_t u = start + blockIdx.x * blockDim.x + threadIdx.x; u < end;
         u += stride) {
        if (parents[u] != -1) continue;
        int32_t par = -1;
        if (bu_scan<1>(g, unstaged(frontier), g.rbeg(u), g.rend(u), 0, 0, par, scanned)) {
            parents[u] = par;
            atomicOr(&next[u >> 5], 1u << (u & 31));
            nnew++;
        }
    }
    for (int s = 32; s > 0; s >>= 1) nnew += __shfl_xor(nnew, s, 64);
    if (lane_id() == 0 && nnew) atomicAdd(num_nodes, (int32_t)nnew);
}

template <typename OffT>
__global__ void bfs_pull_wave_kernel(GraphView<OffT> g,
                                     int32_t *__restrict__ parents,
                                     int32_t start, int32_t end,
                                     const uint32_t *__restrict__ frontier,
                                     uint32_t *__restrict__ next,
                                     int32_t *num_nodes) {
    uint32_t nnew = 0, scanned = 0;
    const int32_t waves = gridDim.x * (blockDim.x / kWave);
    const int32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    for (int32_t u = start + wid; u < end; u += waves) {
        if (parents[u] != -1) continue;
        int32_t par = -1;
        if (bu_scan<64>(g, unstaged(frontier), g.rbeg(u), g.rend(u), lane_id(), 0, par,
                        scanned) &&
            lane_id() == 0) {
            parents[u] = par;
            atomicOr(&next[u >> 5], 1u << (u & 31));
            nnew++;
        }
    }
    if (lane_id() == 0 && nnew) atomicAdd(num_nodes, (int32_t)nnew);
}

// Top-down epoch of the epoch API: wavefront per frontier vertex,
// compare-and-swap on parents, queue output, scout count.
template <typename OffT>
__global__ void bfs_push_epoch_kernel(GraphView<OffT> g,
                                      const int32_t *__restrict__ frontier,
                                      int32_t fsize, int32_t *__restrict__ parents,
                                      int32_t *__restrict__ next_q,
                                      int32_t *next_size, long long *num_edges) {
    const int32_t waves = gridDim.x * (blockDim.x / kWave);
    const int32_t wid = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    unsigned long long eacc = 0;
    for (int32_t f = wid; f < fsize; f += waves) {
        const int32_t u = frontier[f];
        const OffT b = g.rbeg(u), e = g.rend(u);
        for (OffT i = b + OffT(lane_id()); i < e; i += OffT(kWave)) {
            int32_t v = g.adj[i];
            bool won = parents[v] == -1 && atomicCAS(&parents[v], -1, u) == -1;
            if (won) eacc += (unsigned long long)(g.rend(v) - g.rbeg(v));
            wave_append(next_q, reinterpret_cast<uint32_t *>(next_size), won, v);
        }
    }
    for (int s = 32; s > 0; s >>= 1) eacc += __shfl_xor(eacc, s, 64);
    if (lane_id() == 0 && eacc)
        atomicAdd(reinterpret_cast<unsigned long long *>(num_edges), eacc);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// first[u] = the first neighbour of u in CSR order, or -1 for an empty row.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void first_neighbour_kernel(GraphView<OffT> g,
                                                                 int32_t *__restrict__ first) {
    for (int u = blockIdx.x * kBlock + threadIdx.x; u < g.n; u += gridDim.x * kBlock) {
        const OffT b = g.rbeg(u), e = g.rend(u);
        first[u] = b < e ? g.adj[b] : -1;
    }
}

struct BfsWs {
    FrontierWs f;
    int32_t *depth = nullptr;  // used when the caller passes no depth array
    int32_t *first = nullptr;  // first neighbour of each vertex (-1: no edges)
    // Unvisited-list segments: off[grid + 1] then len[2][grid] (sized for
    // f.grid workgroups), laid out for rem_plan at rem_grid workgroups.
    int32_t *rem = nullptr;
    const PlanHost *rem_plan = nullptr;
    int rem_grid = 0;
};

// Workgroup b of a grid-strided bottom-up sweep takes units b, b + grid, ...;
// its list segment holds as many slots as those units have vertices.
static int bfs_rem_layout(BfsWs *ws, const PlanHost *ph, int grid, hipStream_t s) {
    if (ws->rem_plan == ph && ws->rem_grid == grid) return HC_OK;
    std::vector<int32_t> off(grid + 1, 0);
    const PlanDev &p = ph->dev;
    for (int c = 2; c < kNumClasses; c++) {
        const int per = kClassPerUnit[c];
        for (int u = p.unit_start[c]; u < p.unit_start[c + 1]; u++) {
            const int first = (u - p.unit_start[c]) * per;
            off[u % grid + 1] += std::min(per, p.cnt[c] - first);
        }
    }
    for (int b = 0; b < grid; b++) off[b + 1] += off[b];
    HC_HIP(hipMemcpyAsync(ws->rem, off.data(), sizeof(int32_t) * (grid + 1),
                          hipMemcpyHostToDevice, s));
    HC_HIP(hipStreamSynchronize(s));
    ws->rem_plan = ph;
    ws->rem_grid = grid;
    return HC_OK;
}

void free_bfs_ws(hc_graph_s *g) {
    if (!g->ws_bfs) return;
    auto *ws = static_cast<BfsWs *>(g->ws_bfs);
    frontier_ws_free(&ws->f);
    if (ws->depth) hipFree(ws->depth);
    if (ws->rem) hipFree(ws->rem);
    if (ws->first) hipFree(ws->first);
    delete ws;
    g->ws_bfs = nullptr;
}

static int get_bfs_ws(hc_graph_s *g, BfsWs **out, hipStream_t s) {
    // The partial slots are sized for the launch grid; re-allocate when the
    // tuned grid changes.
    if (g->ws_bfs && static_cast<BfsWs *>(g->ws_bfs)->f.grid != grid_blocks()) free_bfs_ws(g);
    if (!g->ws_bfs) {
        auto *ws = new BfsWs();
        int rc = frontier_ws_alloc(g, &ws->f);
        if (rc != HC_OK) {
            delete ws;
            return rc;
        }
        if (hipMalloc(&ws->depth, sizeof(int32_t) * std::max(g->n, 1)) != hipSuccess ||
            hipMalloc(&ws->rem, sizeof(int32_t) * (3 * size_t(ws->f.grid) + 1)) != hipSuccess ||
            hipMalloc(&ws->first, sizeof(int32_t) * std::max(g->n, 1)) != hipSuccess) {
            frontier_ws_free(&ws->f);
            if (ws->depth) hipFree(ws->depth);
            if (ws->rem) hipFree(ws->rem);
            delete ws;
            return fail(HC_ERR_OOM, "BFS workspace allocation failed");
        }
        if (g->n > 0) {
            const int nb = std::min((g->n + kBlock - 1) / kBlock, 65536);
            if (g->off64)
                first_neighbour_kernel<<<nb, kBlock, 0, s>>>(view<uint64_t>(g), ws->first);
            else
                first_neighbour_kernel<<<nb, kBlock, 0, s>>>(view<uint32_t>(g), ws->first);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
                g->ws_bfs = ws;
                free_bfs_ws(g);
                return fail(HC_ERR_HIP, "first-neighbour array build failed");
            }
        }
        g->ws_bfs = ws;
    }
    *out = static_cast<BfsWs *>(g->ws_bfs);
    return HC_OK;
}

template <typename OffT>
struct BfsLaunch {
    PlanDev plan;
    GraphView<OffT> g;
    int32_t *depth;
    int32_t *parents;
    const int32_t *first;
    FrontierBufs fb;
    StepCtl *ctl;
    RemLists rl;
    int grid;
    hipStream_t s;
    int32_t nstage;
    int32_t stage_min;
    int32_t lds_words;
    int32_t push_cap;
    int first_all;
    KernelTimer *kt;
    // Epoch k: header slot k & 1, counter set k % 3.
    static int launch(void *arg, long long k) {
        auto *a = static_cast<BfsLaunch *>(arg);
        FrontierBufs fb = a->fb;
        fb.part = sum_set(a->fb.part, k);
        StepCtl *cur = a->ctl + (k & 1), *nxt = a->ctl + ((k + 1) & 1);
        timed_launch(a->kt, bfs_step_kernel<OffT>, dim3(a->grid), dim3(kBlock),
                     uint32_t(a->lds_words * sizeof(uint32_t)), a->s, a->plan, a->g, a->depth,
                     a->parents, static_cast<const int32_t *>(a->first), fb, cur, a->rl,
                     a->nstage, a->stage_min, a->push_cap, a->first_all);
        bfs_between_kernel<OffT><<<compact_grid(a->grid), kBlock, 0, a->s>>>(
            a->g, fb, cur, nxt, a->ctl, sum_set(a->fb.part, k + 1), (int)sizeof(OffT));
        if (hipGetLastError() != hipSuccess) return fail(HC_ERR_HIP, "BFS epoch launch failed");
        return HC_OK;
    }
};

template <typename OffT>
static int run_bfs_loop(hc_graph_s *g, int32_t src, int32_t *depth,
                        int32_t *parents, int direction, int alpha, int beta,
                        hc_run_stats_t *stats, hipStream_t s) {
    BfsWs *ws = nullptr;
    HC_TRY(get_bfs_ws(g, &ws, s));
    const PlanHost *ph = nullptr;
    HC_TRY(get_plan(g, 0, g->n, tuning().bu, &ph));
    int32_t plan_rows = 0;
    for (int c = 1; c < kNumClasses; c++) plan_rows += ph->dev.cnt[c];
    if (!depth) depth = ws->depth;
    GraphView<OffT> v = view<OffT>(g);
    const int32_t nstage = std::min<int32_t>(tuning().stage_bu, ws->f.fb.words);
    // Dynamic LDS: the bitmap head of bottom-up epochs or the staged output
    // of top-down epochs, whichever is larger.
    const int32_t push_cap = tuning().push_stage;
    const int32_t lds_words = std::max<int32_t>(nstage, push_cap);
    const int grid = resident_grid(reinterpret_cast<const void *>(&bfs_step_kernel<OffT>),
                                   kBlock, lds_words * sizeof(uint32_t), ws->f.grid);
    if (grid > ws->f.grid) return fail(HC_ERR_HIP, "BFS grid %d above its workspace", grid);
    HC_TRY(bfs_rem_layout(ws, ph, grid, s));
    const int init_grid = std::max(1, std::min(grid, (g->n + kBlock - 1) / kBlock));
    bfs_init_kernel<OffT><<<init_grid, kBlock, 0, s>>>(
        v, depth, parents, ws->f.fb, ws->f.ctl, src, direction, alpha, beta, 0,
        (long long)g->m, plan_rows, grid, tuning().bu_remaining, tuning().push_chunk);
    HC_HIP(hipGetLastError());
    ws->f.kt.reset();
    const RemLists rl{ws->rem, reinterpret_cast<uint32_t *>(ws->rem + grid + 1)};
    BfsLaunch<OffT> L{ph->dev, v, depth, parents, tuning().bu_first ? ws->first : nullptr,
                      ws->f.fb, ws->f.ctl, rl, grid, s, nstage, tuning().bu_stage_min,
                      lds_words, push_cap, tuning().bu_first == 2, &ws->f.kt};
    const long long cap = (long long)g->n + 2;  // levels <= n
    float ms = 0;
    long long launches = 0;
    StepCtl fc;
    HC_TRY(run_epoch_loop(&ws->f, s, cap, &BfsLaunch<OffT>::launch, &L, &ms,
                          &launches, &fc));
    if (!fc.done) return fail(HC_ERR_HIP, "BFS loop did not finish in %lld epochs", launches);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        // Device wall clock from the init kernel to the controller of the
        // last epoch (launches queued past it are not part of the loop).
        stats->ms = double(fc.t_end - fc.t_start) / device_state().clock_khz;
        stats->epochs = fc.epoch;
        stats->push_epochs = fc.push_epochs;
        stats->pull_epochs = fc.pull_epochs;
        stats->launches = (int32_t)(2 * launches);
        stats->reached = (int64_t)fc.reached;
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

int hc_bfs(hc_graph_t g, int32_t source, int32_t *depth, int32_t *parents,
           const hc_bfs_opts_t *opts, hc_run_stats_t *stats,
           hc_stream_t stream) {
    HC_CHECK_ARG(g, "graph is NULL");
    HC_CHECK_ARG(g->row0 == 0 && g->rows == g->n,
                 "full runs need a whole-graph handle");
    HC_CHECK_ARG(source >= 0 && source < g->n, "source %d outside [0,%d)",
                 source, g->n);
    int alpha = opts && opts->alpha > 0 ? opts->alpha : 15;
    int beta = opts && opts->beta > 0 ? opts->beta : 18;
    int direction = opts ? opts->direction : 0;
    HC_CHECK_ARG(direction >= 0 && direction <= 2, "direction must be 0, 1 or 2");
    HC_TRY(ensure_init());
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    return g->off64 ? run_bfs_loop<uint64_t>(g, source, depth, parents, direction, alpha, beta, stats, s)
                    : run_bfs_loop<uint32_t>(g, source, depth, parents, direction, alpha, beta, stats, s);
}

int hc_bfs_pull_epoch(hc_graph_t g, int mapping, int32_t start_id,
                      int32_t end_id, int32_t *parents,
                      const uint32_t *frontier, uint32_t *next_frontier,
                      int32_t *num_nodes, int blocks, int threads,
                      hc_stream_t stream) {
    HC_CHECK_ARG(g && parents && frontier && next_frontier && num_nodes,
                 "NULL argument");
    HC_CHECK_ARG(g->row0 <= start_id && start_id <= end_id &&
                     end_id <= g->row0 + g->rows,
                 "range [%d,%d) outside the rows held [%d,%d)", start_id,
                 end_id, g->row0, g->row0 + g->rows);
    HC_CHECK_ARG(mapping == HC_MAP_THREAD || mapping == HC_MAP_WAVE ||
                     mapping == HC_MAP_BUCKETED,
                 "BFS pull epochs support THREAD, WAVE and BUCKETED mappings");
    HC_CHECK_ARG(blocks >= 0 && threads >= 0 && threads <= 1024, "bad launch shape");
    if (mapping == HC_MAP_WAVE)
        HC_CHECK_ARG(threads % kWave == 0, "thread count must be a multiple of 64");
    HC_TRY(ensure_init());
    if (start_id == end_id) return HC_OK;
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    if (!threads) threads = 256;
    if (!blocks) blocks = grid_blocks();
    auto go = [&](auto tag) -> int {
        using OffT = decltype(tag);
        GraphView<OffT> v = view<OffT>(g);
        if (mapping == HC_MAP_THREAD) {
            bfs_pull_thread_kernel<OffT><<<blocks, threads, 0, s>>>(
                v, parents, start_id, end_id, frontier, next_frontier, num_nodes);
        } else if (mapping == HC_MAP_WAVE) {
            bfs_pull_wave_kernel<OffT><<<blocks, threads, 0, s>>>(
                v, parents, start_id, end_id, frontier, next_frontier, num_nodes);
        } else {
            const PlanHost *ph = nullptr;
            HC_TRY(get_plan(g, start_id, end_id, tuning().bu, &ph));
            if (ph->total_units == 0) return HC_OK;
            int nb = std::max(1, std::min(ph->total_units, grid_blocks()));
            const int32_t ns = std::min<int32_t>(tuning().stage_bu, (g->n + 31) / 32);
            bfs_pull_fused_kernel<OffT><<<nb, kBlock, ns * sizeof(uint32_t), s>>>(
                ph->dev, v, parents, frontier, next_frontier, num_nodes, ns);
        }
        HC_HIP(hipGetLastError());
        return HC_OK;
    };
    return g->off64 ? go(uint64_t(0)) : go(uint32_t(0));
}

int hc_bfs_push_epoch(hc_graph_t g, const int32_t *frontier,
                      int32_t frontier_size, int32_t *parents,
                      int32_t *next_frontier, int32_t *next_size,
                      int64_t *num_edges, hc_stream_t stream) {
    HC_CHECK_ARG(g && parents && next_frontier && next_size && num_edges,
                 "NULL argument");
    HC_CHECK_ARG(frontier_size >= 0, "frontier_size < 0");
    HC_CHECK_ARG(frontier_size == 0 || frontier, "frontier is NULL");
    HC_CHECK_ARG(g->row0 == 0 && g->rows == g->n,
                 "push epochs need a whole-graph handle");
    HC_TRY(ensure_init());
    if (frontier_size == 0) return HC_OK;
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    int blocks = std::max(1, std::min(grid_blocks(), (frontier_size + 3) / 4));
    if (g->off64)
        bfs_push_epoch_kernel<uint64_t><<<blocks, 256, 0, s>>>(
            view<uint64_t>(g), frontier, frontier_size, parents, next_frontier,
            next_size, reinterpret_cast<long long *>(num_edges));
    else
        bfs_push_epoch_kernel<uint32_t><<<blocks, 256, 0, s>>>(
            view<uint32_t>(g), frontier, frontier_size, parents, next_frontier,
            next_size, reinterpret_cast<long long *>(num_edges));
    HC_HIP(hipGetLastError());
    return HC_OK;
}

}  // extern "C"
