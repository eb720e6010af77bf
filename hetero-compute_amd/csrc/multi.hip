// Batched SSSP: up to 64 sources of one graph in one device-driven epoch
// loop, one lane per source.
//
// The reference runs one source at a time (sssp_pull_gpu,
// src/kernels/gpu/sssp.cuh:42-100, called per source by the benchmarks,
// src/benchmarks/gpu_benchmark.cuh:541-594).  Here the 64 distance arrays are
// stored vertex-major, dist[v * 64 + q], so one wavefront handles a vertex for
// all 64 sources and every neighbour row it reads is a 256-byte coalesced
// segment shared by the sources.
//
// Each epoch relaxes a vertex u only against neighbours whose distance
// changed since u last saw them: chg[v] holds one bit per 8-source group
// (bit k: some source of lanes 8k..8k+7 improved v), and a lane reads
// dist[v][q] only when its group's bit is set.  The invariant "for every edge
// (u,v) and source q, dist[u][q] <= dist[v][q] + w, or v is flagged for q's
// group" holds at every epoch boundary, so an epoch that changes nothing
// leaves every source's Bellman-Ford fixpoint: the exact shortest distances,
// bit-identical to a single-source run.
//
// Mapping: degree-bucketed units of one wavefront -- a chunk of a hub's edge
// list (partial minimum combined by atomicMin), one vertex, 4 vertices of
// degree <= 16 (16 lanes each) or 16 vertices of degree <= 4 (4 lanes each).
// The loop is device-driven: step + controller kernels form a hipGraph of
// kEpochsPerGraph epochs that the host relaunches until the controller
// reports convergence.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/hc_batch.h"
#include "device_common.h"

namespace hc {

constexpr int kQ = 64;               // sources per batch (= lanes)
constexpr int kEpochsPerGraph = 4;   // epochs per hipGraph launch

struct MultiCtl {
    int32_t done;
    int32_t epoch;
    int32_t max_epochs;
    int32_t promote;                // this launch promotes waiting vertices
    int32_t *dist;                  // vertex-major [n][64]
    unsigned long long changed[2];  // vertices changed in epoch e (slot e&1)
    unsigned long long bytes;       // algorithmic bytes of the epochs run
    long long t_start, t_end;
    // Bucketed propagation (delta > 0): an improvement to a value below
    // thresh flags the vertex for the next epoch, one at or above it marks
    // the vertex as waiting; when an epoch changes nothing, thresh rises by
    // delta and a promotion launch flags the waiting groups now below it.
    int32_t thresh;
    int32_t delta;
    int32_t next_min;               // smallest waiting value left by a promotion
    int32_t pad;
    unsigned long long waits;       // waiting marks set since the last promotion
    unsigned long long left;        // vertices still waiting after a promotion
    unsigned long long promotions;  // promotion launches run
    // Push epochs: when the vertices flagged for the next epoch have at most
    // push_edges edges together and none more than kPushMaxDeg, the next
    // epoch pushes from them (atomicMin into their neighbours' rows) instead
    // of sweeping every row for flagged neighbours.
    int32_t push;                   // this launch pushes
    int32_t pad2;
    unsigned long long push_edges;  // 0: never push
    unsigned long long pushes;      // push launches run
    unsigned long long fedges[2];   // edges of the vertices flagged in epoch e
    unsigned long long fmaxdeg[2];  // largest degree among them
};

// Largest degree a push epoch walks with one wavefront per vertex.
constexpr unsigned long long kPushMaxDeg = 1024;

// Byte of vertex u in a word-packed flag array.
__device__ __forceinline__ uint32_t flag_of(const uint32_t *f, int32_t u) {
    return (f[u >> 2] >> ((u & 3) * 8)) & 0xFFu;
}

// 8-bit group mask of a 64-lane ballot: bit k set when any lane 8k..8k+7 is.
__device__ __forceinline__ uint32_t group_mask(uint64_t b) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) m |= ((b >> (8 * k)) & 0xFFull) ? (1u << k) : 0u;
    return m;
}

template <typename OffT>
struct MultiArgs {
    PlanDev p;
    GraphView<OffT> g;
    uint32_t *chg[3];
    uint32_t *wait;  // waiting groups per vertex (bytes, as chg)
    MultiCtl *ctl;
    int32_t n;
    int32_t unit;  // 1: every edge weighs 1 (breadth-first depths)
};

template <typename OffT>
__device__ __forceinline__ OffT wave_max(OffT v) {
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) {
        OffT o = __shfl_xor(v, s, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Relaxes V = 64/G vertices against their flagged neighbours.  Lane l works
// on sub-vertex l / G, whose edges are [my_b, my_e) (empty for padding);
// best[k] is this lane's (source's) candidate for sub-vertex k.  Returns a
// bit per k: some neighbour of sub-vertex k was flagged for some group.
template <int G, typename OffT>
__device__ __forceinline__ uint32_t relax_group(const MultiArgs<OffT> &a,
                                                const uint32_t *cur,
                                                const int32_t *dist, OffT my_b,
                                                OffT my_e, int32_t *best,
                                                unsigned long long &bytes) {
    constexpr int V = kQ / G;
    const int lane = lane_id();
    const int li = lane % G;
    const int grp = lane >> 3;
    const OffT maxlen = G == kQ ? my_e - my_b : wave_max<OffT>(my_e - my_b);
    uint32_t touched = 0;
    for (OffT it = 0; it < maxlen; it += OffT(G)) {
        const OffT i = my_b + it + OffT(li);
        int32_t v = 0;
        uint32_t m = 0;
        int32_t w = 0;
        const bool in = i < my_e;
        if (in) {
            v = ldnt(&a.g.adj[i]);
            m = flag_of(cur, v);
            if (m) w = a.unit ? 1 : a.g.wt[i];
        }
        uint64_t bal = __ballot(m != 0);
        bytes += 5ull * (unsigned long long)__popcll(__ballot(in));
        while (bal) {
            // Up to four flagged neighbours per round, their rows loaded
            // before any is used.
            int j[4];
            int32_t vj[4], wj[4];
            uint32_t mj[4];
            int32_t dv[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (bal) {
                    j[r] = __ffsll((unsigned long long)bal) - 1;
                    bal &= bal - 1;
                    vj[r] = __builtin_amdgcn_readlane(v, j[r]);
                    wj[r] = __builtin_amdgcn_readlane(w, j[r]);
                    mj[r] = __builtin_amdgcn_readlane(m, j[r]);
                } else {
                    j[r] = -1;
                    vj[r] = 0;
                    wj[r] = 0;
                    mj[r] = 0;
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                dv[r] = HC_INF_DIST;
                if ((mj[r] >> grp) & 1u) dv[r] = dist[(size_t)vj[r] * kQ + lane];
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (j[r] < 0) continue;
                const int sv = j[r] / G;  // wave-uniform
                const int32_t c = relax(dv[r], wj[r]);
#pragma unroll
                for (int k = 0; k < V; k++)
                    if (k == sv) best[k] = min(best[k], c);
                touched |= 1u << sv;
                bytes += 4ull + 32ull * (unsigned long long)__popc(mj[r]);
            }
        }
    }
    return touched;
}

// Per-wavefront counts of one launch, summed per workgroup at its end.
struct EpochAcc {
    unsigned long long bytes = 0;
    unsigned long long fedges = 0;  // edges of the vertices flagged for the next epoch
    unsigned long long fmax = 0;    // their largest degree
    int changed = 0;                // vertices flagged for the next epoch
    int waits = 0;                  // vertices marked waiting (promotion: left waiting)
    int32_t lo = HC_INF_DIST;       // promotion: smallest value left waiting
};

// Flags the groups of vertex u (degree deg) whose sources improved (imp):
// groups with a new value below thresh in nxt, the others in wait.
__device__ __forceinline__ void flag_improved(uint32_t *nxt, uint32_t *wait, int32_t u,
                                              unsigned long long deg, bool imp, int32_t value,
                                              int32_t thresh, EpochAcc &acc) {
    const uint64_t ib = __ballot(imp);
    if (!ib) return;
    const uint64_t now = __ballot(imp && value < thresh);
    const uint64_t later = ib & ~now;
    const int sh = (u & 3) * 8;
    if (now) {
        if (lane_id() == 0) atomicOr(&nxt[u >> 2], group_mask(now) << sh);
        acc.changed++;
        acc.fedges += deg;
        acc.fmax = deg > acc.fmax ? deg : acc.fmax;
    }
    if (later) {
        if (lane_id() == 0) atomicOr(&wait[u >> 2], group_mask(later) << sh);
        acc.waits++;
    }
}

// Commits best[k] for the V sub-vertices (vertex of sub-vertex k is lane
// k*G's my_u, its degree lane k*G's my_deg) and flags improved groups.
// Plain row update: the unit owns these vertices.
template <int V>
__device__ __forceinline__ void commit(int32_t *dist, uint32_t *nxt, uint32_t *wait,
                                       int32_t thresh, int32_t my_u, int32_t my_deg,
                                       const int32_t *best, uint32_t touched, EpochAcc &acc) {
    constexpr int G = kQ / V;
    const int lane = lane_id();
#pragma unroll
    for (int k = 0; k < V; k++) {
        if (!((touched >> k) & 1u)) continue;  // wave-uniform
        const int32_t u = __builtin_amdgcn_readlane(my_u, k * G);
        const int32_t deg = __builtin_amdgcn_readlane(my_deg, k * G);
        int32_t *row = dist + (size_t)u * kQ;
        const int32_t old = row[lane];
        const bool imp = best[k] < old;
        if (imp) row[lane] = best[k];
        acc.bytes += 256ull + 4ull * (unsigned long long)__popcll(__ballot(imp));
        flag_improved(nxt, wait, u, (unsigned long long)deg, imp, best[k], thresh, acc);
    }
}

// Promotion launch: every waiting group of a vertex whose smallest value in
// the group is now below thresh is flagged for the next epoch and stops
// waiting; the others keep waiting (acc.waits counts their vertices, acc.lo
// is their smallest value).  A wavefront owns a whole word of 4 vertices.
template <typename OffT>
__device__ __forceinline__ void promote_waiting(const MultiArgs<OffT> &a, uint32_t *nxt,
                                                const int32_t *dist, int32_t thresh,
                                                int wave0, int wstride, EpochAcc &acc) {
    const int lane = lane_id();
    const int words = (a.n + 3) >> 2;
    for (int w = wave0; w < words; w += wstride) {
        const uint32_t word = a.wait[w];
        acc.bytes += 4;
        if (!word) continue;  // wave-uniform
        uint32_t keep_w = 0, now_w = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const uint32_t m = (word >> (8 * b)) & 0xFFu;
            if (!m) continue;
            const int32_t u = 4 * w + b;
            int32_t x = dist[(size_t)u * kQ + lane];
            acc.bytes += 256;
            // smallest value of each 8-lane group
#pragma unroll
            for (int s = 1; s < 8; s <<= 1) x = min(x, __shfl_xor(x, s, kWave));
            const uint32_t now = group_mask(__ballot(x < thresh)) & m;
            const uint32_t keep = m & ~now;
            if ((keep >> (lane >> 3)) & 1u) acc.lo = min(acc.lo, x);
            keep_w |= keep << (8 * b);
            now_w |= now << (8 * b);
            if (now) {
                const unsigned long long deg =
                    (unsigned long long)(a.g.rend(u) - a.g.rbeg(u));
                acc.changed++;
                acc.fedges += deg;
                acc.fmax = deg > acc.fmax ? deg : acc.fmax;
            }
            acc.waits += keep != 0;
        }
        if (lane == 0) {
            a.wait[w] = keep_w;
            if (now_w) nxt[w] = now_w;  // cleared by the launch before; only writer
        }
    }
}

// Push launch: every vertex v flagged in cur lowers, for each of its flagged
// groups' sources, its neighbours' rows to dist[v] + w by atomicMin (lane q
// = source q, four edges in flight), and flags the neighbours it improved.
// A wavefront owns a word of 4 flagged vertices; the controller chose this
// launch only when those vertices have few edges, none above kPushMaxDeg.
template <typename OffT>
__device__ __forceinline__ void push_flagged(const MultiArgs<OffT> &a, const uint32_t *cur,
                                             uint32_t *nxt, int32_t *dist, int32_t thresh,
                                             int wave0, int wstride, EpochAcc &acc) {
    const int lane = lane_id();
    const int words = (a.n + 3) >> 2;
    for (int w = wave0; w < words; w += wstride) {
        const uint32_t word = cur[w];
        acc.bytes += 4;
        if (!word) continue;  // wave-uniform
#pragma unroll 1
        for (int b = 0; b < 4; b++) {
            const uint32_t m = (word >> (8 * b)) & 0xFFu;
            if (!m) continue;
            const int32_t v = 4 * w + b;
            const int32_t dv = dist[(size_t)v * kQ + lane];
            const bool mine = ((m >> (lane >> 3)) & 1u) && dv != HC_INF_DIST;
            const OffT eb = a.g.rbeg(v), ee = a.g.rend(v);
            acc.bytes += 256ull + 8ull * (unsigned long long)(ee - eb) * (1ull + 8ull * __popc(m));
            for (OffT e = eb; e < ee; e += OffT(4)) {
                int32_t u[4], c[4], old[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const bool ok = e + OffT(r) < ee;  // wave-uniform
                    u[r] = ok ? a.g.adj[e + r] : -1;
                    const int32_t wt = ok ? (a.unit ? 1 : a.g.wt[e + r]) : 0;
                    c[r] = relax(dv, wt);
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    old[r] = HC_INF_DIST;
                    if (u[r] >= 0 && mine && c[r] < HC_INF_DIST)
                        old[r] = atomicMin(&dist[(size_t)u[r] * kQ + lane], c[r]);
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (u[r] < 0) continue;  // wave-uniform
                    const bool imp = mine && c[r] < old[r];
                    if (!__ballot(imp)) continue;
                    const unsigned long long deg =
                        (unsigned long long)(a.g.rend(u[r]) - a.g.rbeg(u[r]));
                    flag_improved(nxt, a.wait, u[r], deg, imp, c[r], thresh, acc);
                }
            }
        }
    }
}

// BLOCK threads per workgroup (knob multi.block): the sweep works in
// wavefront units, so the workgroup size only sets how many share a CU slot.
template <typename OffT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void multi_step_kernel(MultiArgs<OffT> a) {
    MultiCtl *ctl = a.ctl;
    if (ctl->done) return;
    const int ep = ctl->epoch;
    int32_t *dist = ctl->dist;
    const uint32_t *cur = a.chg[ep % 3];
    uint32_t *nxt = a.chg[(ep + 1) % 3];
    uint32_t *clr = a.chg[(ep + 2) % 3];
    const int words = (a.n + 3) >> 2;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < words; i += gridDim.x * BLOCK)
        clr[i] = 0;

    const PlanDev &p = a.p;
    const int lane = lane_id();
    const int32_t thresh = ctl->thresh;
    const int wstride = gridDim.x * (BLOCK / kWave);
    const int wave0 = blockIdx.x * (BLOCK / kWave) + threadIdx.x / kWave;
    EpochAcc acc;
    if (ctl->promote) {
        promote_waiting(a, nxt, dist, thresh, wave0, wstride, acc);
    } else if (ctl->push) {
        push_flagged(a, cur, nxt, dist, thresh, wave0, wstride, acc);
    } else {
        // Wave units: chunks, then vertices of classes 1-2, then groups of 4
        // (class 3), then groups of 16 (classes 4-5).
        const int n0 = p.cnt[0];
        const int n12 = p.cnt[1] + p.cnt[2];
        const int n3 = (p.cnt[3] + 3) / 4;
        const int n45 = (p.cnt[4] + p.cnt[5] + 15) / 16;
        const int total = n0 + n12 + n3 + n45;
        auto vid = [&](int c, int i) -> int32_t {
            return p.list[c] ? p.list[c][i] : p.beg[c] + i;
        };
        for (int unit = wave0; unit < total; unit += wstride) {
            if (unit < n0) {
                const HubChunk hc = p.chunks[unit];
                int32_t best[1] = {HC_INF_DIST};
                uint32_t t =
                    relax_group<kQ>(a, cur, dist, OffT(hc.b), OffT(hc.e), best, acc.bytes);
                if (t) {
                    const int32_t old = atomicMin(&dist[(size_t)hc.u * kQ + lane], best[0]);
                    acc.bytes += 256ull;
                    // a hub: its degree keeps the next epoch a sweep
                    flag_improved(nxt, a.wait, hc.u, ~0ull >> 1, best[0] < old, best[0], thresh,
                                  acc);
                }
            } else if (unit < n0 + n12) {
                const int i = unit - n0;
                const int32_t u = i < p.cnt[1] ? vid(1, i) : vid(2, i - p.cnt[1]);
                const OffT b = a.g.rbeg(u), e = a.g.rend(u);
                int32_t best[1] = {HC_INF_DIST};
                uint32_t t = relax_group<kQ>(a, cur, dist, b, e, best, acc.bytes);
                commit<1>(dist, nxt, a.wait, thresh, u, (int32_t)min(e - b, OffT(INT32_MAX)), best,
                          t, acc);
            } else if (unit < n0 + n12 + n3) {
                const int i = (unit - n0 - n12) * 4 + lane / 16;
                const bool ok = i < p.cnt[3];
                const int32_t u = ok ? vid(3, i) : 0;
                const OffT b = ok ? a.g.rbeg(u) : OffT(0), e = ok ? a.g.rend(u) : OffT(0);
                int32_t best[4] = {HC_INF_DIST, HC_INF_DIST, HC_INF_DIST, HC_INF_DIST};
                uint32_t t = relax_group<16>(a, cur, dist, b, e, best, acc.bytes);
                commit<4>(dist, nxt, a.wait, thresh, u, (int32_t)(e - b), best, t, acc);
            } else {
                const int c4 = p.cnt[4];
                const int i = (unit - n0 - n12 - n3) * 16 + lane / 4;
                const bool ok = i < c4 + p.cnt[5];
                const int32_t u = ok ? (i < c4 ? vid(4, i) : vid(5, i - c4)) : 0;
                const OffT b = ok ? a.g.rbeg(u) : OffT(0), e = ok ? a.g.rend(u) : OffT(0);
                int32_t best[16];
#pragma unroll
                for (int k = 0; k < 16; k++) best[k] = HC_INF_DIST;
                uint32_t t = relax_group<4>(a, cur, dist, b, e, best, acc.bytes);
                commit<16>(dist, nxt, a.wait, thresh, u, (int32_t)(e - b), best, t, acc);
            }
        }
    }
    // One set of atomics per workgroup.  In a promotion launch acc.waits
    // counts the vertices left waiting and acc.lo is their smallest value.
    __shared__ unsigned long long s_bytes, s_changed, s_waits, s_fedges, s_fmax;
    __shared__ int32_t s_lo;
    if (threadIdx.x == 0) {
        s_bytes = 0;
        s_changed = 0;
        s_waits = 0;
        s_fedges = 0;
        s_fmax = 0;
        s_lo = HC_INF_DIST;
    }
    __syncthreads();
    int32_t lo = acc.lo;
#pragma unroll
    for (int s = kWave / 2; s > 0; s >>= 1) lo = min(lo, __shfl_xor(lo, s, kWave));
    if (lane == 0) {
        atomicAdd(&s_bytes, acc.bytes);
        if (acc.changed) {
            atomicAdd(&s_changed, (unsigned long long)acc.changed);
            atomicAdd(&s_fedges, acc.fedges);
            atomicMax(&s_fmax, acc.fmax);
        }
        if (acc.waits) atomicAdd(&s_waits, (unsigned long long)acc.waits);
        if (lo < HC_INF_DIST) atomicMin(&s_lo, lo);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&ctl->bytes, s_bytes);
        if (s_changed) {
            atomicAdd(&ctl->changed[ep & 1], s_changed);
            atomicAdd(&ctl->fedges[ep & 1], s_fedges);
            atomicMax(&ctl->fmaxdeg[ep & 1], s_fmax);
        }
        if (s_waits) atomicAdd(ctl->promote ? &ctl->left : &ctl->waits, s_waits);
        if (s_lo < HC_INF_DIST) atomicMin(&ctl->next_min, s_lo);
    }
}

__global__ void multi_ctl_kernel(MultiCtl *ctl, unsigned long long sweep_bytes) {
    if (threadIdx.x != 0 || ctl->done) return;
    const int ep = ctl->epoch;
    const unsigned long long c = ctl->changed[ep & 1];
    const unsigned long long fe = ctl->fedges[ep & 1], fm = ctl->fmaxdeg[ep & 1];
    ctl->changed[(ep + 1) & 1] = 0;
    ctl->fedges[(ep + 1) & 1] = 0;
    ctl->fmaxdeg[(ep + 1) & 1] = 0;
    ctl->bytes += sweep_bytes;  // offsets and the flag clear, per epoch
    ctl->epoch = ep + 1;
    ctl->t_end = wall_clock64();
    if (ep + 1 >= ctl->max_epochs) {
        ctl->done = 1;
        return;
    }
    const bool promoted = ctl->promote != 0;
    ctl->promote = 0;
    // The next epoch relaxes the flagged vertices: by pushing from them when
    // they are few and light, else by a sweep over every row.
    ctl->push = c != 0 && fe <= ctl->push_edges && fm <= kPushMaxDeg;
    if (ctl->push) ctl->pushes++;
    if (c != 0) return;
    // Nothing flagged: raise the threshold past the waiting values, or stop
    // when none wait.  After a promotion that flagged nothing, the next
    // bucket starts at the smallest value still waiting.
    const unsigned long long waiting = promoted ? ctl->left : ctl->left + ctl->waits;
    if (waiting == 0) {
        ctl->done = 1;
        return;
    }
    const long long base = promoted ? (long long)ctl->next_min : (long long)ctl->thresh;
    ctl->thresh = (int32_t)min(base + (long long)ctl->delta, (long long)HC_INF_DIST);
    ctl->promote = 1;
    ctl->promotions++;
    ctl->waits = 0;
    ctl->left = 0;
    ctl->next_min = HC_INF_DIST;
}

// Fills dist with HC_INF_DIST and clears the flag arrays (3) and the
// waiting array, stored back to back: 4 x words words.
__global__ void multi_fill_kernel(int32_t *dist, size_t count, uint32_t *flags, int words) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += stride)
        dist[i] = HC_INF_DIST;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < 4 * (size_t)words;
         i += stride)
        flags[i] = 0;
}

struct MultiSources {
    int32_t v[kQ];
};

__global__ void multi_seed_kernel(MultiCtl *ctl, int32_t *dist, uint32_t *f0,
                                  MultiSources src, int num, int max_epochs, int delta,
                                  unsigned long long push_edges, int first_push) {
    const int q = threadIdx.x;
    if (q < num) {
        const int32_t s = src.v[q];
        dist[(size_t)s * kQ + q] = 0;
        atomicOr(&f0[s >> 2], (1u << (q >> 3)) << ((s & 3) * 8));
    }
    if (q == 0) {
        ctl->done = max_epochs <= 0 ? 1 : 0;
        ctl->epoch = 0;
        ctl->max_epochs = max_epochs;
        ctl->dist = dist;
        ctl->changed[0] = 0;
        ctl->changed[1] = 0;
        ctl->bytes = 0;
        ctl->promote = 0;
        ctl->delta = delta;
        ctl->thresh = delta > 0 ? delta : HC_INF_DIST;
        ctl->next_min = HC_INF_DIST;
        ctl->waits = 0;
        ctl->left = 0;
        ctl->promotions = 0;
        // The sources are flagged for epoch 0: push from them when allowed.
        ctl->push = first_push;
        ctl->push_edges = push_edges;
        ctl->pushes = first_push;
        ctl->fedges[0] = ctl->fedges[1] = 0;
        ctl->fmaxdeg[0] = ctl->fmaxdeg[1] = 0;
        ctl->t_start = wall_clock64();
        ctl->t_end = ctl->t_start;
    }
}

// Per-graph workspace: flags, controller, and the captured epoch graph.
struct MultiWs {
    uint32_t *flags = nullptr;  // 3 flag arrays and the waiting array, `words` words each
    MultiCtl *ctl = nullptr;
    MultiCtl *h_ctl = nullptr;  // pinned mirror for polling
    int words = 0;
    int grid = 0;
    int block = 0;  // threads per workgroup of the step kernel
    const PlanHost *plan = nullptr;
    // Epoch graphs, [0] weighted, [1] unit weights.
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    KernelTimer kt;
    ~MultiWs() {
        for (int k = 0; k < 2; k++) {
            if (exec[k]) hipGraphExecDestroy(exec[k]);
            if (graph[k]) hipGraphDestroy(graph[k]);
        }
        if (flags) hipFree(flags);
        if (ctl) hipFree(ctl);
        if (h_ctl) hipHostFree(h_ctl);
    }
};

void free_multi_ws(hc_graph_s *g) {
    delete static_cast<MultiWs *>(g->ws_multi);
    g->ws_multi = nullptr;
}

static int get_multi_ws(hc_graph_s *g, MultiWs **out) {
    auto *ws = static_cast<MultiWs *>(g->ws_multi);
    const int grid = grid_blocks(tuning().multi_grid_per_cu);
    const int block = tuning().multi_block;
    if (ws && (ws->grid != grid || ws->block != block)) {  // epoch graphs bake both in
        free_multi_ws(g);
        ws = nullptr;
    }
    if (!ws) {
        ws = new (std::nothrow) MultiWs();
        if (!ws) return fail(HC_ERR_OOM, "host allocation failed");
        ws->words = (g->n + 3) / 4;
        ws->grid = grid;
        ws->block = block;
        hipError_t e1 = hipMalloc(&ws->flags, sizeof(uint32_t) * 4 * (size_t)std::max(ws->words, 1));
        hipError_t e2 = hipMalloc(&ws->ctl, sizeof(MultiCtl));
        hipError_t e3 = hipHostMalloc(&ws->h_ctl, sizeof(MultiCtl), hipHostMallocDefault);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            delete ws;
            return fail(HC_ERR_OOM, "batched-SSSP workspace allocation failed");
        }
        g->ws_multi = ws;
    }
    *out = ws;
    return HC_OK;
}

template <typename OffT>
static const void *multi_step_fn(int block) {
    switch (block) {
        case 128: return reinterpret_cast<const void *>(&multi_step_kernel<OffT, 128>);
        case 512: return reinterpret_cast<const void *>(&multi_step_kernel<OffT, 512>);
        case 1024: return reinterpret_cast<const void *>(&multi_step_kernel<OffT, 1024>);
        default: return reinterpret_cast<const void *>(&multi_step_kernel<OffT, 256>);
    }
}

template <typename OffT>
static int launch_multi_epochs(const MultiArgs<OffT> &a, int grid, int block, int epochs,
                               unsigned long long sweep, hipStream_t s, KernelTimer *kt) {
    const void *step = multi_step_fn<OffT>(block);
    MultiArgs<OffT> args = a;
    void *kargs[] = {(void *)&args};
    for (int i = 0; i < epochs; i++) {
        if (kt) kt->mark(s);
        if (hipLaunchKernel(step, dim3(grid), dim3(block), kargs, 0, s) != hipSuccess)
            return fail(HC_ERR_HIP, "batched SSSP launch failed");
        if (kt) kt->mark(s);
        multi_ctl_kernel<<<1, kWave, 0, s>>>(a.ctl, sweep);
    }
    return hipGetLastError() == hipSuccess ? HC_OK
                                           : fail(HC_ERR_HIP, "batched SSSP launch failed");
}

// Bucket width of the batched SSSP sweep (knob multi.delta): 12 x the
// sampled largest weight / mean degree on flat graphs (uniform 23, weights
// 1..255, degree 32: 96, measured best among 32..128), off on skewed ones
// (Kronecker 24: every width tried was slower).
static int multi_delta(hc_graph_s *g) {
    const int d = tuning().multi_delta;
    if (d >= 0) return d;
    if (!g->weighted || g->m == 0 || g->max_weight <= 0 || !flat_graph(g)) return 0;
    const double mean_deg = (double)g->m / std::max(1, g->rows);
    return (int)std::max(1.0, std::ceil(12.0 * g->max_weight / mean_deg));
}

template <typename OffT>
static int run_multi(hc_graph_s *g, const int32_t *sources, int num, int32_t *dist,
                     int max_epochs, int unit, hc_run_stats_t *stats, hipStream_t s) {
    MultiWs *ws = nullptr;
    HC_TRY(get_multi_ws(g, &ws));
    // Degree classes of the batched sweep (tunable): hub chunks, wavefront,
    // 16-lane and 4-lane vertices.
    const PlanHost *ph = nullptr;
    HC_TRY(get_plan(g, 0, g->n, tuning().multi, &ph));
    if (ph != ws->plan) {  // epoch graphs bake the plan in
        for (int k = 0; k < 2; k++) {
            if (ws->exec[k]) hipGraphExecDestroy(ws->exec[k]);
            if (ws->graph[k]) hipGraphDestroy(ws->graph[k]);
            ws->exec[k] = nullptr;
            ws->graph[k] = nullptr;
        }
        ws->plan = ph;
    }
    uint32_t *f[4] = {ws->flags, ws->flags + ws->words, ws->flags + 2 * (size_t)ws->words,
                      ws->flags + 3 * (size_t)ws->words};
    MultiArgs<OffT> a{ws->plan->dev, view<OffT>(g), {f[0], f[1], f[2]}, f[3], ws->ctl, g->n,
                      unit};
    const int delta = unit ? 0 : multi_delta(g);
    const int grid = ws->grid;
    const unsigned long long sweep =
        (unsigned long long)sizeof(OffT) * (g->n + 1ull) + 4ull * ws->words;
    MultiSources src{};
    for (int q = 0; q < num; q++) src.v[q] = sources[q];
    const size_t cells = (size_t)g->n * kQ;
    const int fill_grid = std::max(1, std::min(grid, (int)((cells + kBlock - 1) / kBlock)));
    multi_fill_kernel<<<fill_grid, kBlock, 0, s>>>(dist, cells, ws->flags, ws->words);
    // Without a limit from the caller the controller runs to convergence
    // (Bellman-Ford: at most n + 1 epochs; with buckets the host's cap below
    // bounds the loop) and the host reports a loop that hits its cap.
    // Push epochs while the flagged vertices hold at most m / multi.push_div
    // edges; epoch 0 pushes from the sources under the same rule.
    const int push_div = tuning().multi_push_div;
    const unsigned long long push_edges =
        push_div > 0 ? std::max<unsigned long long>(1, (unsigned long long)g->m / push_div) : 0;
    unsigned long long src_edges = 0, src_max = 0;
    for (int q = 0; q < num; q++) {
        const int64_t r = sources[q] - g->row0;
        const unsigned long long d =
            (r >= 0 && r < g->rows) ? (unsigned long long)(g->h_off[r + 1] - g->h_off[r]) : 0;
        src_edges += d;
        src_max = std::max(src_max, d);
    }
    const bool first_push = push_edges > 0 && src_edges <= push_edges && src_max <= kPushMaxDeg;
    multi_seed_kernel<<<1, kQ, 0, s>>>(ws->ctl, dist, f[0], src, num,
                                       max_epochs > 0 ? max_epochs : INT32_MAX, delta,
                                       push_edges, first_push ? 1 : 0);
    HC_HIP(hipGetLastError());

    ws->kt.reset();
    const bool use_graph = tuning().multi_graph != 0 && !ws->kt.on;
    if (use_graph && !ws->exec[unit]) {
        // The kernels read dist and the epoch from the controller, so one
        // graph serves every call on this graph handle.  It is built node
        // by node rather than captured from a stream: capture would forbid
        // other threads (serving slots) from touching the legacy stream
        // meanwhile.
        hipGraph_t graph = nullptr;
        HC_HIP(hipGraphCreate(&graph, 0));
        ws->graph[unit] = graph;
        hipGraphNode_t prev = nullptr;
        for (int i = 0; i < kEpochsPerGraph; i++) {
            hipKernelNodeParams kp{};
            void *step_args[] = {(void *)&a};
            kp.func = const_cast<void *>(multi_step_fn<OffT>(ws->block));
            kp.gridDim = dim3(grid);
            kp.blockDim = dim3(ws->block);
            kp.sharedMemBytes = 0;
            kp.kernelParams = step_args;
            kp.extra = nullptr;
            hipGraphNode_t step = nullptr, ctl = nullptr;
            HC_HIP(hipGraphAddKernelNode(&step, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp));
            MultiCtl *cp = a.ctl;
            unsigned long long sw = sweep;
            void *ctl_args[] = {(void *)&cp, (void *)&sw};
            hipKernelNodeParams cpar{};
            cpar.func = reinterpret_cast<void *>(&multi_ctl_kernel);
            cpar.gridDim = dim3(1);
            cpar.blockDim = dim3(kWave);
            cpar.sharedMemBytes = 0;
            cpar.kernelParams = ctl_args;
            cpar.extra = nullptr;
            HC_HIP(hipGraphAddKernelNode(&ctl, graph, &step, 1, &cpar));
            prev = ctl;
        }
        HC_HIP(hipGraphInstantiate(&ws->exec[unit], graph, nullptr, nullptr, 0));
    }
    // Relaunch until the controller reports convergence; the done flag of
    // batch b-1 is read while batch b runs.
    hipEvent_t evb[2];
    HC_HIP(hipEventCreateWithFlags(&evb[0], hipEventDisableTiming));
    HC_HIP(hipEventCreateWithFlags(&evb[1], hipEventDisableTiming));
    MultiCtl *h = ws->h_ctl;
    h->done = 0;
    const long long bound = max_epochs > 0 ? max_epochs
                            : delta > 0    ? 64LL * (g->n + 1) + 4096
                                           : (long long)g->n + 1;
    const long long cap = bound + 2 * kEpochsPerGraph;
    long long launched = 0;
    bool done = false;
    int rc = HC_OK;
    for (int b = 0; !done && launched < cap; b++) {
        if (use_graph) {
            if (hipGraphLaunch(ws->exec[unit], s) != hipSuccess) {
                rc = fail(HC_ERR_HIP, "graph launch failed");
                break;
            }
        } else {
            rc = launch_multi_epochs(a, grid, ws->block, kEpochsPerGraph, sweep, s, &ws->kt);
            if (rc != HC_OK) break;
        }
        launched += kEpochsPerGraph;
        hipMemcpyAsync(&h->done, &ws->ctl->done, sizeof(int32_t), hipMemcpyDeviceToHost, s);
        hipEventRecord(evb[b & 1], s);
        if (b > 0) {
            hipEventSynchronize(evb[(b - 1) & 1]);
            done = __atomic_load_n(&h->done, __ATOMIC_ACQUIRE) != 0;
        }
    }
    hipEventDestroy(evb[0]);
    hipEventDestroy(evb[1]);
    HC_TRY(rc);
    MultiCtl c;
    HC_HIP(hipMemcpyAsync(&c, ws->ctl, sizeof(MultiCtl), hipMemcpyDeviceToHost, s));
    HC_HIP(hipStreamSynchronize(s));
    if (!c.done) return fail(HC_ERR_HIP, "batched SSSP did not converge in %lld epochs", launched);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = double(c.t_end - c.t_start) / device_state().clock_khz;
        stats->epochs = c.epoch;
        const int32_t pushes = (int32_t)c.pushes;
        stats->push_epochs = pushes;
        stats->pull_epochs = c.epoch - (int32_t)c.promotions - pushes;
        stats->launches = (int32_t)(2 * launched + 2);
        stats->reached = num;
        stats->alg_bytes = (double)c.bytes;
        if (ws->kt.on) {
            stats->kernel_launches = std::min(c.epoch, ws->kt.pairs());
            stats->kernel_ms = ws->kt.sum_ms(stats->kernel_launches);
        }
    }
    return HC_OK;
}

}  // namespace hc

using namespace hc;

__global__ void multi_extract_kernel(const int32_t *__restrict__ mat, int32_t n, int col,
                                     int32_t *__restrict__ out, int minus1) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < (size_t)n; v += stride) {
        const int32_t x = mat[v * kQ + col];
        out[v] = (minus1 && x == HC_INF_DIST) ? -1 : x;
    }
}

// Order-independent checksum of one column of a vertex-major matrix (stride
// 64) or of a plain array (stride 1): the 64-bit wrapping sum over v of
// mix(v, value), value -1 for HC_INF_DIST when minus1 is set (BFS depths
// of a batch, so that they match a single-source run's).
__global__ void checksum_kernel(const int32_t *__restrict__ data, int32_t n, int stride,
                                int col, int minus1, unsigned long long *out) {
    unsigned long long h = 0;
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < (size_t)n; v += step) {
        int32_t x = data[v * (size_t)stride + col];
        if (minus1 && x == HC_INF_DIST) x = -1;
        h += checksum_mix((uint64_t)v, x);
    }
    const unsigned long long t = block_sum_u64(h);
    if (threadIdx.x == 0) atomicAdd(out, t);
}

// The same sums for every column of a vertex-major [n][64] matrix in one
// pass: lane l of each wave reads column l of kSumRows consecutive rows
// (256 contiguous bytes a row), the waves' sums meet in LDS and one atomic
// add per column and workgroup lands in out.
constexpr int kSumRows = 4;

__global__ void checksum_cols_kernel(const int32_t *__restrict__ m, int32_t n, int ncols,
                                     int minus1, unsigned long long *out) {
    __shared__ unsigned long long part[kWavesPerBlock][kQ];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    unsigned long long h = 0;
    const size_t step = (size_t)gridDim.x * kWavesPerBlock * kSumRows;
    for (size_t r0 = ((size_t)blockIdx.x * kWavesPerBlock + w) * kSumRows; r0 < (size_t)n;
         r0 += step) {
        int32_t x[kSumRows];
#pragma unroll
        for (int i = 0; i < kSumRows; i++)
            x[i] = r0 + i < (size_t)n ? m[(r0 + i) * kQ + lane] : 0;
#pragma unroll
        for (int i = 0; i < kSumRows; i++) {
            if (r0 + i >= (size_t)n) break;
            const int32_t v = minus1 && x[i] == HC_INF_DIST ? -1 : x[i];
            h += checksum_mix((uint64_t)(r0 + i), v);
        }
    }
    part[w][lane] = h;
    __syncthreads();
    if (threadIdx.x < ncols) {
        unsigned long long t = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; i++) t += part[i][threadIdx.x];
        atomicAdd(&out[threadIdx.x], t);
    }
}

__global__ void zero_words_kernel(unsigned long long *p, int count) {
    if ((int)threadIdx.x < count) p[threadIdx.x] = 0;
}

// Device accumulators of the checksum entry points: 64 words per call,
// taken from a per-device free list (allocated on first need, kept until
// the process ends), so concurrent calls on different streams do not wait
// for each other.
struct SumWords {
    unsigned long long *p = nullptr;
    int dev = -1;
    static std::mutex mu;
    static std::vector<unsigned long long *> free_list[64];

    int take() {
        HC_HIP(hipGetDevice(&dev));
        HC_CHECK_ARG(dev >= 0 && dev < 64, "device %d above the 64 supported", dev);
        {
            std::lock_guard<std::mutex> lk(mu);
            if (!free_list[dev].empty()) {
                p = free_list[dev].back();
                free_list[dev].pop_back();
                return HC_OK;
            }
        }
        HC_HIP(hipMalloc(reinterpret_cast<void **>(&p), kQ * sizeof(unsigned long long)));
        return HC_OK;
    }
    // Copies count sums to the host and waits for them.
    int fetch(uint64_t *dst, int count, hipStream_t s) {
        HC_HIP(hipMemcpyAsync(dst, p, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, s));
        HC_HIP(hipStreamSynchronize(s));
        return HC_OK;
    }
    ~SumWords() {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu);
        free_list[dev].push_back(p);
    }
};
std::mutex SumWords::mu;
std::vector<unsigned long long *> SumWords::free_list[64];

static int multi_entry(hc_graph_t g, const int32_t *sources, int32_t num_sources,
                       int32_t *dist, int32_t max_epochs, int unit, hc_run_stats_t *stats,
                       hc_stream_t stream) {
    HC_CHECK_ARG(g && sources && dist, "NULL argument");
    HC_CHECK_ARG(unit || g->weighted, "SSSP needs a weighted graph");
    HC_CHECK_ARG(g->row0 == 0 && g->rows == g->n, "batched runs need a whole-graph handle");
    HC_CHECK_ARG(num_sources >= 1 && num_sources <= kQ, "num_sources must be in [1, %d]", kQ);
    for (int q = 0; q < num_sources; q++)
        HC_CHECK_ARG(sources[q] >= 0 && sources[q] < g->n, "source %d = %d outside [0,%d)", q,
                     sources[q], g->n);
    HC_TRY(ensure_init());
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    if (unit && tuning().msbfs)
        return run_msbfs_entry(g, sources, num_sources, dist, max_epochs, stats, s);
    return g->off64
               ? run_multi<uint64_t>(g, sources, num_sources, dist, max_epochs, unit, stats, s)
               : run_multi<uint32_t>(g, sources, num_sources, dist, max_epochs, unit, stats, s);
}

extern "C" {

int hc_sssp_multi(hc_graph_t g, const int32_t *sources, int32_t num_sources, int32_t *dist,
                  int32_t max_epochs, hc_run_stats_t *stats, hc_stream_t stream) {
    return multi_entry(g, sources, num_sources, dist, max_epochs, 0, stats, stream);
}

int hc_bfs_multi(hc_graph_t g, const int32_t *sources, int32_t num_sources, int32_t *depth,
                 int32_t max_epochs, hc_run_stats_t *stats, hc_stream_t stream) {
    return multi_entry(g, sources, num_sources, depth, max_epochs, 1, stats, stream);
}

int hc_checksum_i32(const int32_t *data, int32_t num_nodes, int32_t stride, int32_t column,
                    int32_t unreached_minus1, uint64_t *sum, hc_stream_t stream) {
    HC_CHECK_ARG(data && sum, "NULL argument");
    HC_CHECK_ARG(num_nodes >= 0 && (stride == 1 || stride == kQ), "stride must be 1 or %d", kQ);
    HC_CHECK_ARG(column >= 0 && column < stride, "bad column %d", column);
    HC_TRY(ensure_init());
    *sum = 0;
    if (num_nodes == 0) return HC_OK;
    SumWords acc;
    HC_TRY(acc.take());
    hipStream_t s = pick_stream(stream);
    const int grid = std::max(1, std::min(grid_blocks(), (num_nodes + kBlock - 1) / kBlock));
    zero_words_kernel<<<1, kQ, 0, s>>>(acc.p, 1);
    checksum_kernel<<<grid, kBlock, 0, s>>>(data, num_nodes, stride, column, unreached_minus1,
                                            acc.p);
    HC_HIP(hipGetLastError());
    return acc.fetch(sum, 1, s);
}

int hc_checksum_multi(const int32_t *matrix, int32_t num_nodes, int32_t num_columns,
                      int32_t unreached_minus1, uint64_t *sums, hc_stream_t stream) {
    HC_CHECK_ARG(matrix && sums, "NULL argument");
    HC_CHECK_ARG(num_nodes >= 0 && num_columns >= 1 && num_columns <= kQ,
                 "num_columns must be in 1..%d", kQ);
    HC_TRY(ensure_init());
    for (int c = 0; c < num_columns; c++) sums[c] = 0;
    if (num_nodes == 0) return HC_OK;
    SumWords acc;
    HC_TRY(acc.take());
    hipStream_t s = pick_stream(stream);
    const int rows_per_block = kWavesPerBlock * kSumRows;
    const int grid =
        std::max(1, std::min(grid_blocks(), (num_nodes + rows_per_block - 1) / rows_per_block));
    zero_words_kernel<<<1, kQ, 0, s>>>(acc.p, num_columns);
    checksum_cols_kernel<<<grid, kBlock, 0, s>>>(matrix, num_nodes, num_columns,
                                                 unreached_minus1, acc.p);
    HC_HIP(hipGetLastError());
    return acc.fetch(sums, num_columns, s);
}

int hc_multi_extract(const int32_t *matrix, int32_t num_nodes, int32_t column,
                     int32_t unreached_minus1, int32_t *out, hc_stream_t stream) {
    HC_CHECK_ARG(matrix && out, "NULL argument");
    HC_CHECK_ARG(num_nodes >= 0 && column >= 0 && column < kQ, "bad column %d", column);
    HC_TRY(ensure_init());
    if (num_nodes == 0) return HC_OK;
    hipStream_t s = pick_stream(stream);
    const int grid = std::max(1, std::min(grid_blocks(), (num_nodes + kBlock - 1) / kBlock));
    multi_extract_kernel<<<grid, kBlock, 0, s>>>(matrix, num_nodes, column, out,
                                                 unreached_minus1);
    HC_HIP(hipGetLastError());
    return HC_OK;
}

}  // extern "C"
