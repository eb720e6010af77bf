// Device-resident frontier and epoch-control state shared by the
// device-driven epoch loops (SSSP push/pull and BFS top-down/bottom-up).
//
// Every epoch is two launches on one stream:
//   step      (full grid)  runs the epoch; each workgroup adds its counts
//                          (frontier size, degree sum, edges examined) into
//                          one of kSumGroups counter groups;
//   between   (compaction grid) every workgroup sums the counters and
//                          applies the controller's decision to its own LDS
//                          copy of the control header (the same inputs give
//                          the same result everywhere); workgroup 0 records
//                          the epoch and stores the next header; after a
//                          pull epoch followed by a push epoch, the grid
//                          turns the frontier bitmap into the degree-classed
//                          queues the push expansion reads.
// The control header is double-buffered (epoch k reads slot k & 1, its
// controller writes slot (k + 1) & 1) and the counter groups are
// triple-buffered (the step of epoch k adds into set k % 3, the controller
// of epoch k clears set (k + 1) % 3), so no launch writes what another
// workgroup of the same launch reads.  The host passes k with each launch.
// All cross-workgroup data flows through kernel boundaries, so no device-
// scope fences are needed inside a kernel, and the host only enqueues.
//
// Push frontiers are routed by degree into classes that map to different
// lane counts when expanded:
//   chunk: deg > kHubDeg      -> one workgroup per chunk of StepCtl::chunk edges
//   q[0] : deg > kBigDeg      -> one workgroup per vertex
//   q[1] : deg > kMidDeg      -> one wavefront per vertex
//   q[2] : otherwise          -> kSmallLanes lanes per vertex
// Chunks are small because a push epoch is latency-bound (each edge is a
// dependent gather + atomic): many short chunks keep all CUs busy when the
// frontier holds only a few hubs.
#ifndef HC_FRONTIER_H
#define HC_FRONTIER_H

#include "device_common.h"

namespace hc {

constexpr int64_t kHubDeg = 1024;
constexpr int64_t kMinChunkEdges = 128;  // smallest push chunk (sizes the chunk lists)
constexpr int64_t kBigDeg = 256;
constexpr int64_t kMidDeg = 16;
constexpr int kSmallLanes = 8;
constexpr int kSmallPerUnit = kBlock / kSmallLanes;
constexpr int kTraceLen = 128;

enum StepMode : int32_t { kPush = 0, kPull = 1, kPromote = 2 };

struct StepCtl {
    int32_t done;
    int32_t mode;
    int32_t epoch;
    int32_t max_epochs;
    int32_t direction;             // 0 auto, 1 pull only, 2 push only
    int32_t alpha;
    int32_t beta;
    int32_t need_compact;          // next push epoch needs queues from bitmap
    int32_t n;                     // vertices
    int32_t grid;                  // workgroups of the step launches
    int32_t push_epochs;
    int32_t pull_epochs;
    long long m;                   // edges of the graph
    long long edges_to_check;      // BFS direction-optimizing state
    long long scout;
    long long num_nodes;           // BFS pull phase: current frontier size
    long long prev_nodes;
    long long in_n;                // size of the frontier the next epoch reads
    long long in_e;                // its degree sum
    long long sweep_bytes;         // algorithmic bytes of one pull sweep
    unsigned long long reached;
    unsigned long long alg_bytes;  // algorithmic bytes moved by the epochs
    int32_t plan_rows;
    int32_t filter;                // SSSP: next pull gathers changed neighbours only
    int32_t filter_knob;           // SSSP: filtering allowed (tuning knob)
    int32_t rem_valid;             // BFS: the next pull reads the unvisited lists
    int32_t rem_knob;              // BFS: lanes per listed vertex, 0: no lists
    int32_t chunk;                 // edges per hub chunk of push epochs
    // SSSP buckets (delta > 0, push epochs only): a vertex lowered to a value
    // at or above thresh goes to the far bitmap instead of the next
    // frontier; when a frontier comes out empty the threshold rises (by
    // delta, doubled after a promotion that found nothing) and a promotion
    // epoch moves the far vertices now below it into the frontier.
    int32_t thresh;
    int32_t delta;
    int32_t far_step;              // the last rise of thresh
    int32_t promotions;            // promotion epochs run
    long long far_live;            // upper bound of the far vertices
    // Device wall clock (constant rate): start of the loop, end of the last
    // epoch, and end of each traced epoch.
    long long t_start;
    long long t_end;
    long long tr_t[kTraceLen];
    // Per-epoch trace of the first kTraceLen epochs.
    int8_t tr_mode[kTraceLen];
    uint32_t tr_n[kTraceLen];
    unsigned long long tr_e[kTraceLen];
    unsigned long long tr_scan[kTraceLen];
};

// Step sums.  Every workgroup adds its totals into one of kSumGroups counter
// groups (kSumStride words apart, each in its own cache line), so no counter
// takes more than grid / kSumGroups same-address atomics; the controller
// reads and clears them.  Slots: frontier vertices produced, their degree
// sum, edges examined, and two algorithm-specific counts.
constexpr int kSumGroups = 64;
constexpr int kSumStride = 16;
constexpr int kSumSlots = 5;
static_assert(kSumGroups <= kWave, "one wavefront reads the counter groups");

constexpr int kSumSets = 3;
constexpr int kSumWords = kSumGroups * kSumStride;

struct Partials {
    unsigned long long *acc;   // [kSumGroups][kSumStride]
};

// Counter set j of the kSumSets consecutive sets at base.
__host__ __device__ inline Partials sum_set(const Partials &base, long long j) {
    return Partials{base.acc + (j % kSumSets) * kSumWords};
}

struct FrontierBufs {
    int32_t *q[2][3];
    HubChunk *ch[2];
    uint32_t *bm[3];
    uint32_t *far;             // SSSP buckets: vertices waiting for a higher threshold
    uint32_t (*cnt)[5];        // queue fill [parity][q0, q1, q2, chunks, hubs]
    int32_t words;             // bitmap words
    Partials part;             // counter set of the current epoch
    int32_t *host_done;        // pinned host word the controller sets at the end
};

// Tells the polling host that the loop has finished (system-scope store
// into pinned host memory; no copy on the stream).
__device__ __forceinline__ void publish_done(const FrontierBufs &fb) {
    if (fb.host_done) __hip_atomic_store(fb.host_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Workspace layout for one graph (SSSP and BFS share the shape).
struct FrontierWs {
    FrontierBufs fb{};
    StepCtl *ctl = nullptr;
    StepCtl *h_ctl = nullptr;  // pinned mirror for polling
    void *dmem = nullptr;
    int64_t chunk_cap = 0;
    int32_t grid = 0;
    KernelTimer kt;  // step-kernel launches of the current run
};

int frontier_ws_alloc(hc_graph_s *g, FrontierWs *ws);
void frontier_ws_free(FrontierWs *ws);

__device__ __forceinline__ int degree_class(int64_t d) {
    return d > kHubDeg ? 3 : (d > kBigDeg ? 0 : (d > kMidDeg ? 1 : 2));
}

// Appends v (when pred) to the degree-classed queues of parity p.
// Wavefront-aggregated; used by push epochs, whose output is bounded by the
// direction heuristic, so the per-wavefront atomics stay few.
template <typename OffT>
__device__ __forceinline__ void frontier_append(const GraphView<OffT> &g,
                                                const FrontierBufs &fb,
                                                StepCtl *ctl, int p, bool pred,
                                                int32_t v) {
    int cls = -1;
    OffT b = 0, e = 0;
    if (pred) {
        b = g.rbeg(v);
        e = g.rend(v);
        cls = degree_class(int64_t(e - b));
    }
    wave_append(fb.q[p][0], &fb.cnt[p][0], cls == 0, v);
    wave_append(fb.q[p][1], &fb.cnt[p][1], cls == 1, v);
    wave_append(fb.q[p][2], &fb.cnt[p][2], cls == 2, v);
    // Hub vertices: their chunk slots are reserved with one atomic per
    // wavefront, and the active lanes write each hub's chunk descriptors
    // together (a hub can have hundreds).
    const uint64_t m3 = __ballot(cls == 3);
    if (m3) {
        const int64_t ck = ctl->chunk;
        const uint32_t nch = cls == 3 ? (uint32_t)((int64_t(e - b) + ck - 1) / ck) : 0u;
        uint32_t total = 0;
        for (uint64_t t = m3; t; t &= t - 1) total += __shfl(nch, __ffsll((unsigned long long)t) - 1);
        const int leader = __ffsll((unsigned long long)m3) - 1;
        uint32_t base = 0;
        if (lane_id() == leader) {
            base = atomicAdd(&fb.cnt[p][3], total);
            atomicAdd(&fb.cnt[p][4], (uint32_t)__popcll(m3));
        }
        base = __shfl(base, leader);
        const uint64_t act = __ballot(true);
        const uint32_t rank = (uint32_t)__popcll(act & lanemask_lt());
        const uint32_t nact = (uint32_t)__popcll(act);
        for (uint64_t t = m3; t; t &= t - 1) {
            const int l = __ffsll((unsigned long long)t) - 1;
            const uint32_t n_l = __shfl(nch, l);
            const int32_t v_l = __shfl(v, l);
            const long long b_l = __shfl((long long)b, l);
            const long long e_l = __shfl((long long)e, l);
            for (uint32_t j = rank; j < n_l; j += nact) {
                const long long cb = b_l + (long long)j * ck;
                const long long ce = cb + ck < e_l ? cb + ck : e_l;
                fb.ch[p][base + j] = HubChunk{v_l, 0, cb, ce};
            }
            base += n_l;
        }
    }
}

// Push-epoch output staged per workgroup in LDS.  Appending straight to the
// degree-classed queues costs one atomic per wavefront and class on the same
// few counters, and same-address atomics serialize at about 12 ns each on
// this GPU (a big top-down epoch issued ~10^5 of them); staged output is
// classified at the end of the sweep and reserved with one atomic per class
// and workgroup.  Entries past `cap` go to the queues directly.

// With `dedup` (SSSP: the next-frontier bitmap), staged entries are
// candidates that may repeat (a vertex lowered by several edges of one
// epoch); the flush keeps the entry that sets the vertex's bit first and
// counts it, so the sweep's per-edge chain holds one returning atomic (the
// minimum), not two.  Entries past `cap` are deduplicated on the spot.
struct PushOut {
    int32_t *buf;      // LDS, cap entries
    uint32_t *count;   // LDS: entries appended (past cap: sent to the queues)
    uint32_t cap;
    uint32_t *dedup;   // bitmap that deduplicates candidates, or nullptr
};

// Counts of the entries a deduplicating flush kept (frontier vertices and
// their degree sum).
struct PushKept {
    uint32_t n;
    unsigned long long e;
};

template <typename OffT>
__device__ __forceinline__ void push_out_append(const PushOut &o, const GraphView<OffT> &g,
                                                const FrontierBufs &fb, StepCtl *ctl, int p,
                                                bool pred, int32_t v, uint32_t &kept_n,
                                                unsigned long long &kept_e) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(o.count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    const uint32_t pos = base + (uint32_t)__popcll(m & lanemask_lt());
    const bool fits = pos < o.cap;
    if (pred && fits) o.buf[pos] = v;
    bool direct = pred && !fits;
    if (o.dedup && direct) {
        const uint32_t bit = 1u << (v & 31);
        direct = !(atomicOr(&o.dedup[v >> 5], bit) & bit);
        if (direct) {
            kept_n++;
            kept_e += (unsigned long long)(g.rend(v) - g.rbeg(v));
        }
    }
    frontier_append(g, fb, ctl, p, direct, v);
}

// Moves the staged entries into the queues of parity p (all threads of the
// workgroup; after the sweep, or earlier to make room).  Returns this
// thread's share of the kept counts when deduplicating (zero otherwise).
template <typename OffT>
__device__ PushKept push_out_flush(const PushOut &o, const GraphView<OffT> &g,
                                   const FrontierBufs &fb, StepCtl *ctl, int p) {
    __shared__ uint32_t s_n[5], s_base[5];
    PushKept kept{0u, 0ull};
    __syncthreads();
    const uint32_t n = min(*o.count, o.cap);
    if (threadIdx.x < 5) s_n[threadIdx.x] = 0;
    __syncthreads();
    const int64_t ck = ctl->chunk;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int32_t v = o.buf[i];
        const int64_t d = int64_t(g.rend(v) - g.rbeg(v));
        if (o.dedup) {
            const uint32_t bit = 1u << (v & 31);
            if (atomicOr(&o.dedup[v >> 5], bit) & bit) {
                o.buf[i] = -1;  // a repeat: dropped
                continue;
            }
            kept.n++;
            kept.e += (unsigned long long)d;
        }
        const int c = degree_class(d);
        if (c < 3) {
            atomicAdd(&s_n[c], 1u);
        } else {
            atomicAdd(&s_n[3], (uint32_t)((d + ck - 1) / ck));
            atomicAdd(&s_n[4], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        const uint32_t k = s_n[threadIdx.x];
        s_base[threadIdx.x] = k ? atomicAdd(&fb.cnt[p][threadIdx.x], k) : 0u;
        s_n[threadIdx.x] = 0;  // now the write cursor of the class
    }
    __syncthreads();
    // Every lane takes part in every round (a hub's chunk descriptors, up to
    // hundreds, are written by the whole wavefront).
    for (uint32_t i0 = 0; i0 < n; i0 += blockDim.x) {
        const uint32_t i = i0 + threadIdx.x;
        int32_t v = 0;
        int c = -1;
        OffT b = 0, e = 0;
        if (i < n && o.buf[i] >= 0) {
            v = o.buf[i];
            b = g.rbeg(v);
            e = g.rend(v);
            c = degree_class(int64_t(e - b));
        }
        if (c >= 0 && c < 3) fb.q[p][c][s_base[c] + atomicAdd(&s_n[c], 1u)] = v;
        const uint32_t nch = c == 3 ? (uint32_t)((int64_t(e - b) + ck - 1) / ck) : 0u;
        const uint32_t at = c == 3 ? s_base[3] + atomicAdd(&s_n[3], nch) : 0u;
        for (uint64_t t = __ballot(c == 3); t; t &= t - 1) {
            const int l = __ffsll((unsigned long long)t) - 1;
            const uint32_t n_l = __shfl(nch, l), at_l = __shfl(at, l);
            const int32_t v_l = __shfl(v, l);
            const long long b_l = __shfl((long long)b, l), e_l = __shfl((long long)e, l);
            for (uint32_t j = lane_id(); j < n_l; j += kWave) {
                const long long cb = b_l + (long long)j * ck;
                fb.ch[p][at_l + j] = HubChunk{v_l, 0, cb, cb + ck < e_l ? cb + ck : e_l};
            }
        }
    }
    return kept;
}

// Called by the whole workgroup between units of a push sweep: flushes the
// staged output early when the next unit could overfill the buffer (a unit
// appends at most `unit_max` entries: its edges), so the per-lane queue
// atomics of the overflow path stay a last resort.  Returns the staged count
// afterwards (the same in every thread).  The second barrier keeps every
// thread's read of the count ahead of the next unit's appends, so all
// threads take the same branch.
template <typename OffT>
__device__ __forceinline__ uint32_t push_out_make_room(const PushOut &o, const GraphView<OffT> &g,
                                                       const FrontierBufs &fb, StepCtl *ctl,
                                                       int p, uint32_t unit_max,
                                                       uint32_t &kept_n,
                                                       unsigned long long &kept_e) {
    __syncthreads();
    const uint32_t n = *o.count;
    __syncthreads();
    if (n == 0 || n + unit_max <= o.cap) return n;
    const PushKept k = push_out_flush(o, g, fb, ctl, p);
    kept_n += k.n;
    kept_e += k.e;
    __syncthreads();
    if (threadIdx.x == 0) *o.count = 0;
    __syncthreads();
    return 0;
}

// Entries a push unit can append: a hub chunk's edges, or at most kHubDeg
// (a workgroup row, four wavefront rows of at most kBigDeg, or 32 small rows
// of at most kMidDeg).
__device__ __forceinline__ uint32_t push_unit_max(uint32_t unit, uint32_t nch, int64_t chunk) {
    return unit < nch ? (uint32_t)chunk : (uint32_t)kHubDeg;
}

// Chunk descriptors of a hub source (parity 0), written by one workgroup.
template <typename OffT>
__device__ __forceinline__ void write_source_chunks(const GraphView<OffT> &g,
                                                    const FrontierBufs &fb, int32_t src,
                                                    int64_t chunk) {
    const int64_t b = int64_t(g.rbeg(src)), e = int64_t(g.rend(src));
    if (degree_class(e - b) != 3) return;
    const int64_t nch = (e - b + chunk - 1) / chunk;
    for (int64_t j = threadIdx.x; j < nch; j += blockDim.x) {
        const int64_t cb = b + j * chunk;
        fb.ch[0][j] = HubChunk{src, 0, cb, cb + chunk < e ? cb + chunk : e};
    }
}

// Workgroup sums of three per-thread counters plus two per-workgroup values
// (x, y: taken from thread 0), added into this workgroup's counter group.
__device__ __forceinline__ void write_partials(const Partials &part, uint32_t n,
                                               unsigned long long e,
                                               unsigned long long s,
                                               unsigned long long x = 0,
                                               unsigned long long y = 0) {
    __shared__ uint32_t s_n[kWavesPerBlock];
    __shared__ unsigned long long s_e[kWavesPerBlock], s_s[kWavesPerBlock];
#pragma unroll
    for (int sh = 32; sh > 0; sh >>= 1) {
        n += __shfl_xor(n, sh, 64);
        e += __shfl_xor(e, sh, 64);
        s += __shfl_xor(s, sh, 64);
    }
    __syncthreads();
    if (lane_id() == 0) {
        s_n[threadIdx.x / kWave] = n;
        s_e[threadIdx.x / kWave] = e;
        s_s[threadIdx.x / kWave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long v[kSumSlots] = {0, 0, 0, x, y};
        for (int i = 0; i < kWavesPerBlock; i++) {
            v[0] += s_n[i];
            v[1] += s_e[i];
            v[2] += s_s[i];
        }
        unsigned long long *g = part.acc + (blockIdx.x % kSumGroups) * kSumStride;
        for (int k = 0; k < kSumSlots; k++)
            if (v[k]) atomicAdd(&g[k], v[k]);
    }
}

// Reads the step sums of one counter set; run by every workgroup of a
// launch that reads them (integer sums: exact in any order); results valid
// in every thread.
__device__ __forceinline__ void sum_partials(const Partials &part,
                                             unsigned long long (&out)[kSumSlots]) {
    __shared__ unsigned long long r[kSumSlots];
    if (threadIdx.x < kWave) {
        unsigned long long v[kSumSlots];
        const unsigned long long *g = part.acc + threadIdx.x * kSumStride;
        for (int k = 0; k < kSumSlots; k++) v[k] = threadIdx.x < kSumGroups ? g[k] : 0ull;
        for (int k = 0; k < kSumSlots; k++) {
#pragma unroll
            for (int sh = 32; sh > 0; sh >>= 1) v[k] += __shfl_xor(v[k], sh, 64);
        }
        if (threadIdx.x == 0)
            for (int k = 0; k < kSumSlots; k++) r[k] = v[k];
    }
    __syncthreads();
    for (int k = 0; k < kSumSlots; k++) out[k] = r[k];
}

// Zeroes one counter set (one workgroup).
__device__ __forceinline__ void clear_partials(const Partials &part) {
    for (int i = threadIdx.x; i < kSumWords; i += blockDim.x) part.acc[i] = 0;
}

// Records epoch k: the end time in the controller's copy c, the trace entry
// in `trace` (slot 0 of the control block; null in all workgroups but one).
__device__ __forceinline__ void trace_epoch(StepCtl *c, StepCtl *trace, int k, int mode,
                                            unsigned long long out_n,
                                            unsigned long long out_e,
                                            unsigned long long scanned) {
    const long long t = wall_clock64();
    c->t_end = t;
    if (trace && k < kTraceLen) {
        trace->tr_t[k] = t;
        trace->tr_mode[k] = (int8_t)mode;
        trace->tr_n[k] = (uint32_t)out_n;
        trace->tr_e[k] = out_e;
        trace->tr_scan[k] = scanned;
    }
}

// The controllers work on an LDS copy of the control block's header (every
// field before the trace arrays): one coalesced load and one coalesced store
// instead of a chain of dependent global accesses by one thread.
constexpr int kCtlHeadWords =
    (int)(__builtin_offsetof(StepCtl, tr_t) / sizeof(unsigned long long));
static_assert(__builtin_offsetof(StepCtl, tr_t) % sizeof(unsigned long long) == 0,
              "the control block header is copied in 8-byte words");
static_assert(kCtlHeadWords <= kBlock, "one word per thread");

__device__ __forceinline__ void ctl_head_copy(StepCtl *dst, const StepCtl *src) {
    if (threadIdx.x < kCtlHeadWords)
        reinterpret_cast<unsigned long long *>(dst)[threadIdx.x] =
            reinterpret_cast<const unsigned long long *>(src)[threadIdx.x];
}

// Prologue of a between-epochs launch (every workgroup): an LDS copy of
// the header epoch k's step read (`cur`) and that step's sums.  False when
// the loop has ended.
__device__ __forceinline__ bool ctl_begin(StepCtl *s_c, const StepCtl *cur, const Partials &part,
                                          unsigned long long (&sums)[kSumSlots]) {
    ctl_head_copy(s_c, cur);
    __syncthreads();
    if (s_c->done) return false;
    sum_partials(part, sums);  // ends with a barrier
    return true;
}

// Epilogue of the controller (after a barrier that follows the decision on
// s_c, which advanced its epoch to k + 1; workgroup 0 only): the header of
// epoch k + 1 goes to its slot `nxt`, the queue counters of parity k & 1
// (read by epoch k, appended to by epoch k + 1) and the counter set of
// epoch k + 1 are cleared, and a finished loop is marked in both header
// slots (launches queued past the end read either) and published to the
// host.
__device__ __forceinline__ void ctl_end(const StepCtl *s_c, StepCtl *cur, StepCtl *nxt,
                                        const FrontierBufs &fb, const Partials &next_set) {
    if (blockIdx.x != 0) return;
    ctl_head_copy(nxt, s_c);
    clear_partials(next_set);
    const int used = (s_c->epoch - 1) & 1;
    if (threadIdx.x < 5) fb.cnt[used][threadIdx.x] = 0;
    if (threadIdx.x == 0 && s_c->done) {
        cur->done = 1;
        publish_done(fb);
    }
}

// Bitmap -> degree-classed queues of parity (epoch & 1), run by the
// between-epochs launch with the decided header c.  Each workgroup
// takes a contiguous run of bitmap words, counts its vertices per class in
// LDS and reserves its output with one atomic per class.
// The launch is queued after every epoch but has compaction work only when
// the direction turns to top-down; a smaller grid keeps its idle launches
// cheap (tuning knob compact.grid: workgroups, at most the step grid).
inline int compact_grid(int step_grid) { return std::max(1, std::min(step_grid, tuning().compact_grid)); }

// SSSP with a 16-bit distance mirror (`d16`): before every pull epoch the
// same launch rebuilds the mirror from the int32 distances.
template <typename OffT>
__device__ void compact_body(const GraphView<OffT> &g, const FrontierBufs &fb,
                             const StepCtl *ctl, int three_bitmaps,
                             const int32_t *dist = nullptr, uint16_t *d16 = nullptr) {
    if (ctl->done) return;
    if (d16 && ctl->mode == kPull) {
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < g.n; i += gridDim.x * kBlock) {
            const int32_t d = dist[i];
            d16[i] = d >= HC_INF_DIST ? uint16_t(0xFFFF) : (d >= 0xFFFE ? uint16_t(0xFFFE) : uint16_t(d));
        }
    }
    if (!ctl->need_compact) return;
    const int k = ctl->epoch;
    const int p = k & 1;
    const uint32_t *bm = three_bitmaps ? fb.bm[k % 3] : fb.bm[k & 1];
    const int words = fb.words;
    const int64_t ck = ctl->chunk;
    // Word w belongs to workgroup w % grid, so the hub words at the start
    // of a degree-sorted bitmap spread over many workgroups.
    const int stride = gridDim.x;
    const int rounds = (words - (int)blockIdx.x + stride * kBlock - 1) / (stride * kBlock);
    constexpr uint32_t kHubList = 256;
    __shared__ uint32_t s_cnt[5], s_base[5], s_nhub;
    __shared__ int32_t s_hub_v[kHubList];
    __shared__ uint32_t s_hub_at[kHubList], s_hub_pre[kHubList + 1];
    __shared__ long long s_hub_b[kHubList], s_hub_e[kHubList];
    for (int r = 0; r < rounds; r++) {
        if (threadIdx.x < 5) s_cnt[threadIdx.x] = 0;
        if (threadIdx.x == 0) s_nhub = 0;
        __syncthreads();
        const int w = (int)blockIdx.x + (r * kBlock + (int)threadIdx.x) * stride;
        uint32_t bits = (w < words) ? bm[w] : 0u;
        // pass 1: count per class (class 3 counts chunks; slot 4 hubs)
        uint32_t my[5] = {0, 0, 0, 0, 0};
        for (uint32_t x = bits; x; x &= x - 1) {
            int32_t v = w * 32 + __ffs(x) - 1;
            int64_t d = int64_t(g.rend(v) - g.rbeg(v));
            int c = degree_class(d);
            if (c < 3) {
                my[c]++;
            } else {
                my[3] += (uint32_t)((d + ck - 1) / ck);
                my[4]++;
            }
        }
        uint32_t off[5];
        for (int c = 0; c < 5; c++) off[c] = my[c] ? atomicAdd(&s_cnt[c], my[c]) : 0;
        __syncthreads();
        if (threadIdx.x < 5)
            s_base[threadIdx.x] = s_cnt[threadIdx.x]
                                      ? atomicAdd(&fb.cnt[p][threadIdx.x], s_cnt[threadIdx.x])
                                      : 0;
        __syncthreads();
        // pass 2: queue entries; hubs are listed with their chunk range and
        // their descriptors written by the whole workgroup below (serially
        // by the owning thread only past the list's capacity).
        for (uint32_t x = bits; x; x &= x - 1) {
            int32_t v = w * 32 + __ffs(x) - 1;
            OffT b = g.rbeg(v), e = g.rend(v);
            int64_t d = int64_t(e - b);
            int c = degree_class(d);
            if (c < 3) {
                fb.q[p][c][s_base[c] + off[c]++] = v;
            } else {
                const uint32_t at = s_base[3] + off[3];
                off[3] += (uint32_t)((d + ck - 1) / ck);
                const uint32_t h = atomicAdd(&s_nhub, 1u);
                if (h < kHubList) {
                    s_hub_v[h] = v;
                    s_hub_at[h] = at;
                    s_hub_b[h] = (long long)b;
                    s_hub_e[h] = (long long)e;
                } else {
                    for (int64_t cb = int64_t(b); cb < int64_t(e); cb += ck) {
                        int64_t ce = cb + ck < int64_t(e) ? cb + ck : int64_t(e);
                        fb.ch[p][at + (uint32_t)((cb - int64_t(b)) / ck)] = HubChunk{v, 0, cb, ce};
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t nh = min(s_nhub, kHubList);
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t h = 0; h < nh; h++) {
                s_hub_pre[h] = t;
                t += (uint32_t)((s_hub_e[h] - s_hub_b[h] + ck - 1) / ck);
            }
            s_hub_pre[nh] = t;
        }
        __syncthreads();
        // Chunk j of the listed hubs: hub h with pre[h] <= j < pre[h+1].
        for (uint32_t j = threadIdx.x; j < s_hub_pre[nh]; j += blockDim.x) {
            uint32_t lo = 0, hi = nh;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (s_hub_pre[mid] <= j) lo = mid; else hi = mid;
            }
            const uint32_t q = j - s_hub_pre[lo];
            const long long cb = s_hub_b[lo] + (long long)q * ck;
            const long long ce = cb + ck < s_hub_e[lo] ? cb + ck : s_hub_e[lo];
            fb.ch[p][s_hub_at[lo] + q] = HubChunk{s_hub_v[lo], 0, cb, ce};
        }
        __syncthreads();
    }
}

// Host loop shared by SSSP and BFS: enqueue `launch_epoch(arg, k)` for
// epochs k = 0, 1, ... in batches and poll the done flag of the previous
// batch while the next one is queued.  final_ctl: the last header with the
// trace of slot 0.
int run_epoch_loop(FrontierWs *ws, hipStream_t s, long long cap,
                   int (*launch_epoch)(void *, long long), void *arg, float *ms,
                   long long *launches, StepCtl *final_ctl);

// The newer header of the two slots at d (epoch counts differ by one) with
// the trace of slot 0.
int read_final_ctl(const StepCtl *d, StepCtl *out, hipStream_t s);

}  // namespace hc

#endif
