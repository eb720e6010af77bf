// Device-side building blocks shared by the epoch kernels (hipcc only).
//
// Wavefronts are 64 lanes on gfx950; every group reduction, ballot and
// aggregated append below is written for that width.  Offsets are templated
// on uint32_t / uint64_t so graphs under 2^32 edges use 4-byte row offsets.
#ifndef HC_DEVICE_COMMON_H
#define HC_DEVICE_COMMON_H

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "hc_internal.h"

#define HC_HIP(expr)                                                        \
    do {                                                                    \
        hipError_t hc_e_ = (expr);                                          \
        if (hc_e_ != hipSuccess)                                            \
            return hc::fail(HC_ERR_HIP, "%s failed: %s (%s:%d)", #expr,     \
                            hipGetErrorString(hc_e_), __FILE__, __LINE__); \
    } while (0)

namespace hc {

constexpr int kWave = 64;
constexpr int kBlock = 256;          // threads per workgroup of the fused kernels
constexpr int kWavesPerBlock = kBlock / kWave;

// ------------------------------------------------------------------------
// Graph view passed by value to kernels.
// ------------------------------------------------------------------------
template <typename OffT>
struct GraphView {
    const OffT *off;     // rows+1 offsets, local to the handle's first row
    const int32_t *adj;  // neighbour ids (global)
    const int32_t *wt;   // weights or nullptr
    int32_t row0;        // global id of local row 0
    int32_t n;           // global vertex count
    __device__ __forceinline__ OffT rbeg(int32_t u) const { return off[u - row0]; }
    __device__ __forceinline__ OffT rend(int32_t u) const { return off[u - row0 + 1]; }
};

// ------------------------------------------------------------------------
// Degree-bucketed plan of a vertex range.
// Classes (in unit order): 0 hub chunks (workgroup per edge chunk),
// 1 workgroup per vertex, 2 wavefront per vertex, 3 16-lane group per vertex,
// 4 4-lane group per vertex, 5 thread per vertex.  Vertices of degree 0 are in
// no class (the zero tail of a degree-sorted graph is skipped).
// ------------------------------------------------------------------------
constexpr int kNumClasses = 6;
constexpr int kClassGroup[kNumClasses] = {kBlock, kBlock, 64, 16, 4, 1};
constexpr int kClassPerUnit[kNumClasses] = {1, 1, 4, 16, 64, 256};

struct HubChunk {
    int32_t u;      // global vertex id
    int32_t hub;    // index of the hub vertex within the plan
    int64_t b, e;   // local edge range of this chunk
};

struct PlanDev {
    // Per class: either a contiguous id range [beg, beg+cnt) or a list.
    int32_t beg[kNumClasses];
    int32_t cnt[kNumClasses];           // vertices (class 0: chunks)
    const int32_t *list[kNumClasses];   // nullptr => range
    int32_t unit_start[kNumClasses + 1];
    const HubChunk *chunks;             // class 0 descriptors
    int32_t num_hubs;                   // distinct hub vertices
    const int32_t *hub_chunk_start;     // num_hubs+1 prefix of chunks per hub
    // Degree-0 vertices of the range (not in any unit): range or list.
    int32_t zero_beg;
    int32_t zero_cnt;
    const int32_t *zero_list;
};

struct PlanConfig {
    int64_t hub_threshold = 16384;  // deg > this -> chunks
    int64_t hub_chunk = 4096;       // edges per chunk
    int64_t block_threshold = 1024; // deg > this -> workgroup per vertex
    int64_t wave_threshold = 192;   // deg > this -> wavefront per vertex
    int64_t g16_threshold = 24;     // deg > this -> 16 lanes
    int64_t g4_threshold = 5;       // deg > this -> 4 lanes, else 1
    // Only rows of degree in [min_degree, max_degree] (min_degree > 0 also
    // drops degree-0 rows from the zero list).
    int64_t min_degree = 0;
    int64_t max_degree = INT64_MAX;
    bool operator<(const PlanConfig &o) const {
        return std::tie(hub_threshold, hub_chunk, block_threshold,
                        wave_threshold, g16_threshold, g4_threshold, min_degree, max_degree) <
               std::tie(o.hub_threshold, o.hub_chunk, o.block_threshold,
                        o.wave_threshold, o.g16_threshold, o.g4_threshold, o.min_degree,
                        o.max_degree);
    }
};

// Process-wide tuning table (set through hc_set_tuning by the autotuner or a
// caller; read when plans and workspaces are built).
// Many-graph BFS limits: graphs of up to kManyLdsNodesMax vertices can keep
// their three bitmaps in LDS (48 KiB); at most kManyHeavyCapMax heavy frontier
// vertices are listed per level.
constexpr int kManyLdsNodesMax = 1 << 17;
constexpr int kManyHeavyCapMax = 1024;

struct Tuning {
    int grid_per_cu = 8;   // workgroups per CU of the fused / step launches
    int sssp_alpha = 0;    // SSSP goes pull when frontier edges > m / alpha (0: by
                           // graph shape, 4 skewed / 2 flat)
    int sssp_grid_per_cu = 4;  // workgroups per CU of the SSSP loop (at most)
    int sssp_dist16 = 2;   // SSSP pulls gather a 16-bit mirror of the distances
                           // (0 off, 1 on, 2: on for flat graphs)
    int sssp_filter = 0;   // SSSP pulls gather only neighbours changed last epoch
    int sssp_push_dedup = 2;  // SSSP pushes deduplicate improved vertices when staging
                              // them (1) or on the spot (0); 2: by graph shape
    int sssp_stage = 0;    // per-vertex entries staged in LDS by SSSP pull sweeps
    int64_t sssp_block_threshold = 4096;  // SSSP: degree above which a WG takes a row
    int64_t sssp_wave_threshold = 96;     // SSSP: ... a wavefront takes a row
    int64_t sssp_g16_threshold = 0;       // SSSP: ... 16 lanes take a row (0: 24)
    int64_t sssp_g4_threshold = 8;        // SSSP: ... 4 lanes take a row (else 1)
    int stage_pull = 8192; // per-vertex entries staged in LDS by pull sweeps
    int stage_bu = 8192;   // frontier-bitmap words staged by bottom-up sweeps
    int push_stage = 8192; // push epochs: output entries staged in LDS per workgroup
    int push_chunk = 256;  // edges per hub chunk of push epochs (>= 128)
    int bu_remaining = 1;  // BFS: lanes per vertex of unvisited-list sweeps (0: off)
    int bu_first = 2;      // BFS: rows probe the first-neighbour array (1: one-lane
                           // rows, 2: all rows)
    int bu_stage_min = 4096;  // BFS: unvisited lists shorter than this skip LDS staging
    int compact_grid = 4096;  // workgroups of the frontier compaction launch (capped)
    int loop_batch = 2;    // epochs the host enqueues per poll of the done flag
    int kernel_events = 0; // 1: time every launch of a loop's main kernel
    int multi_graph = 1;   // batched SSSP/BFS: relaunch an epoch hipGraph
    int multi_grid_per_cu = 16;  // workgroups per CU of the batched sweep
    int multi_block = 256;       // threads per workgroup of the batched sweep (128-1024)
    int sssp_delta = -1;         // single-source SSSP: bucket width (push epochs only;
                                 // 0 off; -1: 2 x max weight / mean degree on flat
                                 // graphs, off on skewed ones)
    int multi_push_div = 16;     // batched sweep: an epoch pushes from the flagged
                                 // vertices when they hold <= m / this edges (0: never)
    int multi_delta = -1;        // batched SSSP: improvements at or above a threshold
                                 // that rises in steps of this wait (delta-stepping
                                 // buckets; 0: off; -1: 12 x max weight / mean degree
                                 // on flat graphs, off on skewed ones)
    int msbfs = 1;               // batched BFS: bit-parallel levels (1) or the batched
                                 // sweep with unit weights (0)
    int msbfs_alpha = 32;        // bit-parallel BFS pushes a level while its frontier's
                                 // edges < m / this, else pulls
    int msbfs_col_shift = -1;    // its pulls gather column blocks of 2^this vertices
                                 // (0: off; -1: four blocks on flat graphs, sixteen
                                 // on skewed ones, at least 2^16 vertices each)
    int many_block = 512;        // many-graph BFS: threads per workgroup (256, 512, 1024)
    int many_per_cu = 2;         // ... workgroups per CU (0: as many as the device holds)
    int many_light_deg = 32;     // ... top-down: a thread expands vertices up to this degree
    int many_heavy_cap = 512;    // ... heavier ones, up to this many per level, by wavefronts
    int many_lds_nodes = 1 << 17;  // ... graphs up to this many vertices keep bitmaps in LDS
    int pull_col_shift = -1;     // pull sweeps gather column blocks of 2^this vertices
                                 // (0: off; -1: automatic, flat graphs only; used
                                 // when the graph has >= 2 blocks)
    int64_t cb_min_slice = 4;    // rows of degree > this x blocks are swept by column
                                 // block, lighter ones once over the plain CSR
    int64_t cb_g4_slice = 2;     // column-blocked sweeps: mean slice degree (row
    int64_t cb_g16_slice = 12;   // degree / blocks) above which 4 / 16 lanes, a
    int64_t cb_wave_slice = 48;  // wavefront or a workgroup take a row's slice
    int64_t cb_block_slice = 512;
    PlanConfig pull;       // SSSP and PageRank pull sweeps
    PlanConfig bu;         // BFS bottom-up sweeps (no chunk / WG classes)
    PlanConfig multi;      // batched sweeps: chunks, wave, 16-lane, 4-lane
    Tuning() {
        bu.hub_threshold = INT64_MAX;
        bu.block_threshold = INT64_MAX;
        bu.wave_threshold = 4096;
        bu.g16_threshold = 512;
        bu.g4_threshold = 64;
        multi.hub_threshold = 2048;
        multi.hub_chunk = 1024;
        multi.block_threshold = INT64_MAX;
        multi.wave_threshold = 32;
        multi.g16_threshold = 4;
        multi.g4_threshold = 0;
    }
};
// The knobs in force on the calling thread: those of the graph whose call
// it is running (a GraphCall), else the process-wide set.
Tuning &tuning();
Tuning &global_tuning();
// Plan of the SSSP pull sweeps: the shared pull hub classes with SSSP's own
// workgroup, wavefront and lane-group thresholds.
inline PlanConfig sssp_plan_config() {
    PlanConfig c = tuning().pull;
    c.wave_threshold = tuning().sssp_wave_threshold;
    const int64_t g16 = tuning().sssp_g16_threshold > 0 ? tuning().sssp_g16_threshold : 24;
    c.g16_threshold = std::min<int64_t>(g16, c.wave_threshold);
    c.g4_threshold = std::min<int64_t>(tuning().sssp_g4_threshold, c.g16_threshold);
    c.block_threshold = std::max<int64_t>(tuning().sssp_block_threshold, c.wave_threshold);
    if (c.hub_threshold < c.block_threshold) c.hub_threshold = c.block_threshold;
    return c;
}

// Column-blocked pull sweeps over K blocks take the rows of degree above
// cb_light_degree(K) block by block (plan cb_plan_config: classes by the
// mean slice degree, row degree / K, no hub chunks); lighter rows, whose
// slices would be mostly empty, are swept once over the plain CSR (plan
// cb_light_config: the plain SSSP classes restricted to them).
inline int64_t cb_light_degree(int K) { return tuning().cb_min_slice * K; }
inline PlanConfig cb_plan_config(int K) {
    const Tuning &t = tuning();
    PlanConfig c;
    c.hub_threshold = INT64_MAX;
    c.hub_chunk = int64_t(1) << 40;
    c.g4_threshold = t.cb_g4_slice * K;
    c.g16_threshold = std::max(c.g4_threshold, t.cb_g16_slice * K);
    c.wave_threshold = std::max(c.g16_threshold, t.cb_wave_slice * K);
    c.block_threshold = std::max(c.wave_threshold, t.cb_block_slice * K);
    c.min_degree = cb_light_degree(K) + 1;
    return c;
}
inline PlanConfig cb_light_config(int K) {
    PlanConfig c = sssp_plan_config();
    c.min_degree = 1;
    c.max_degree = cb_light_degree(K);
    return c;
}

// Per-launch timing of a loop's main kernel with HIP events recorded on the
// kernel's own stream just before and after it (enabled by the
// kernel_events knob).  Events are pooled and reused across runs.
struct KernelTimer {
    std::vector<hipEvent_t> ev;
    int used = 0;
    bool on = false;
    void reset() {
        on = tuning().kernel_events != 0;
        used = 0;
    }
    // Two fresh events for the next timed launch (nullptr when off).
    hipEvent_t *next_pair() {
        if (!on) return nullptr;
        while ((int)ev.size() < used + 2) {
            hipEvent_t e;
            // Device-scope release: the timestamp is not delayed by a
            // system-scope cache writeback after the kernel.
            if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) {
                on = false;
                return nullptr;
            }
            ev.push_back(e);
        }
        used += 2;
        return &ev[used - 2];
    }
    void mark(hipStream_t s) {
        if (!on) return;
        if (used == (int)ev.size()) {
            hipEvent_t e;
            if (hipEventCreateWithFlags(&e, hipEventReleaseToDevice) != hipSuccess) {
                on = false;
                return;
            }
            ev.push_back(e);
        }
        hipEventRecord(ev[used++], s);
    }
    // Sum over the first n (before, after) pairs; the stream must have
    // completed them.
    double sum_ms(int n) const {
        double t = 0;
        for (int i = 0; i < n && 2 * i + 1 < used; i++) {
            float x = 0;
            if (hipEventElapsedTime(&x, ev[2 * i], ev[2 * i + 1]) == hipSuccess) t += x;
        }
        return t;
    }
    int pairs() const { return used / 2; }
    ~KernelTimer() {
        for (hipEvent_t e : ev) hipEventDestroy(e);
    }
};

// Launches `kernel` on s.  With the timer on, the dispatch itself records
// its start and stop events (hipExtLaunchKernelGGL), so the time between
// them is the kernel's execution only: events recorded around a launch
// would also count the wait for the host to enqueue it when the stream has
// run dry.
template <typename K, typename... Args>
inline void timed_launch(KernelTimer *kt, K kernel, dim3 grid, dim3 block, uint32_t lds,
                         hipStream_t s, Args... args) {
    hipEvent_t *e = kt ? kt->next_pair() : nullptr;
    if (e)
        hipExtLaunchKernelGGL(kernel, grid, block, lds, s, e[0], e[1], 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}

// Host-side owner of a plan's device arrays.
struct PlanHost {
    PlanDev dev{};
    void *dmem = nullptr;  // single allocation holding lists and chunks
    int32_t total_units = 0;
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Streaming load of the adjacency / weight arrays: each element is read once
// per sweep, so it is marked non-temporal to leave L2 to the reused
// per-vertex state (distances, scores, frontier bits).
template <typename T>
__device__ __forceinline__ T ldnt(const T *p) {
    return __builtin_nontemporal_load(p);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    return (uint64_t(1) << lane_id()) - 1;
}

template <int G>
__device__ __forceinline__ int32_t group_min(int32_t v) {
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) v = min(v, __shfl_xor(v, s, G));
    return v;
}

template <int G>
__device__ __forceinline__ int64_t group_sum64(int64_t v) {
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) v += __shfl_xor(v, s, G);
    return v;
}

// Saturating distance candidate: d <= HC_INF_DIST and w < 2^31, so the
// unsigned sum cannot wrap; results are clamped to the sentinel.
__device__ __forceinline__ int32_t relax(int32_t d, int32_t w) {
    uint32_t c = uint32_t(d) + uint32_t(w);
    return int32_t(c < uint32_t(HC_INF_DIST) ? c : uint32_t(HC_INF_DIST));
}

// Wavefront-aggregated append: every active lane with pred set gets a
// distinct slot of q; one atomic per wavefront.
__device__ __forceinline__ void wave_append(int32_t *q, uint32_t *count,
                                            bool pred, int32_t value) {
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (pred) q[base + __popcll(m & lanemask_lt())] = value;
}

__device__ __forceinline__ uint32_t load_acquire_u32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t load_acquire_u64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// Reader of a per-vertex array whose first k entries are staged in LDS.  In a
// degree-sorted graph the low ids are the hubs, which receive a large share
// of all neighbour references (about a third of them for the top 16K
// vertices of a Kronecker graph), so those gathers turn into LDS reads and
// the random L1->L2 request stream shrinks accordingly.
template <typename T>
struct Staged {
    const T *g;  // global array
    const T *s;  // LDS copy of g[0, k)
    int32_t k;
    __device__ __forceinline__ T operator[](int32_t v) const { return v < k ? s[v] : g[v]; }
};

template <typename T>
__device__ __forceinline__ Staged<T> unstaged(const T *g) {
    return Staged<T>{g, nullptr, 0};
}

// Copies g[0, k) into the workgroup's LDS buffer (all threads; ends with a
// barrier).  The copy is a snapshot: later writes to g[0, k) by other
// workgroups are not seen through it, which the callers allow for.
template <typename T>
__device__ __forceinline__ Staged<T> stage(const T *g, int32_t k, T *lds) {
    for (int i = threadIdx.x; i < k; i += blockDim.x) lds[i] = g[i];
    __syncthreads();
    return Staged<T>{g, lds, k};
}

// Largest dynamic LDS a staged kernel asks for (leaves room for the static
// LDS of the reductions within a 64 KiB workgroup allocation).
constexpr int kMaxStageBytes = 60 * 1024;

// Sets bit u of bm for every lane with pred; lanes that share a word are
// merged into one atomicOr (the degree-sorted classes hand a wavefront
// consecutive vertices).  Returns whether this lane's bit was newly set.
__device__ __forceinline__ bool wave_set_bit(uint32_t *bm, int32_t u, bool pred) {
    const uint32_t word = pred ? uint32_t(u) >> 5 : 0u;
    const uint32_t bit = pred ? (1u << (u & 31)) : 0u;
    bool first = false;
    uint64_t todo = __ballot(pred);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t w = __shfl(word, leader);
        const bool mine = pred && word == w;
        const uint64_t grp = __ballot(mine);
        uint32_t bits = mine ? bit : 0u;
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) bits |= __shfl_xor(bits, s, 64);
        uint32_t old = 0;
        if (lane_id() == leader) old = atomicOr(&bm[w], bits);
        old = __shfl(old, leader);
        if (mine) first = !(old & bit);
        todo &= ~grp;
    }
    return first;
}

// Sets bit u of bm for every lane with pred, lanes that share a word merged
// into one atomicOr whose old value is not waited for.  For callers whose
// lanes own their vertices in this launch (the bit is new exactly when pred).
__device__ __forceinline__ void wave_or_bits(uint32_t *bm, int32_t u, bool pred) {
    const uint32_t word = pred ? uint32_t(u) >> 5 : 0u;
    const uint32_t bit = pred ? (1u << (u & 31)) : 0u;
    uint64_t todo = __ballot(pred);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t w = __shfl(word, leader);
        const bool mine = pred && word == w;
        const uint64_t grp = __ballot(mine);
        uint32_t bits = mine ? bit : 0u;
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) bits |= __shfl_xor(bits, s, 64);
        if (lane_id() == leader) atomicOr(&bm[w], bits);
        todo &= ~grp;
    }
}

// Sum of a per-thread value over a 256-thread workgroup; valid in thread 0.
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
    __shared__ unsigned long long s_part[kWavesPerBlock];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    __syncthreads();
    if (lane_id() == 0) s_part[threadIdx.x / kWave] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWavesPerBlock; i++) t += s_part[i];
    return t;
}

// ------------------------------------------------------------------------
// Runtime state (host)
// ------------------------------------------------------------------------
struct DeviceState {
    int device = -1;
    int cus = 0;
    double clock_khz = 100000.0;  // device wall clock rate
    size_t hbm = 0;
    hipStream_t stream = nullptr;  // library default stream
    bool ready = false;
};
DeviceState &device_state();
int ensure_init();
// Workgroups for a full-device launch: per_cu per CU (0 = the tuned value).
int grid_blocks(int per_cu = 0);
// Workgroups of `kernel` (block threads, lds_bytes dynamic LDS) that fit on
// the device at once, at most cap: a grid-stride launch of this size has no
// second wave of late-starting workgroups.
int resident_grid(const void *kernel, int block, size_t lds_bytes, int cap);
hipStream_t pick_stream(hc_stream_t s);

}  // namespace hc

// Host-side graph handle.
namespace hc {
// Column-blocked copy of a whole graph's CSR for pull sweeps: the edges of
// row u whose neighbour lies in block k = v >> shift form the slice
// [off[k][u], off[k][u+1]) of adj / wt, in CSR order.  A sweep that takes the
// blocks one after another gathers per-vertex state from one block of 2^shift
// vertices at a time, which one XCD's L2 holds.
struct ColBlocks {
    int32_t shift = 0;
    int32_t K = 0;             // blocks: ceil(n / 2^shift)
    void *off = nullptr;       // [K][n+1] offsets of the graph's offset type
    int32_t *adj = nullptr;
    int32_t *wt = nullptr;     // weighted graphs only
    uint32_t *seen = nullptr;  // ceil(n/32) words: rows counted by an epoch-API sweep
    void *dmem = nullptr;
    size_t bytes = 0;
};
// Knobs set on one graph (and every share of it): key/value pairs applied
// over the process-wide set for the duration of each call on the graph.
struct GraphTuning {
    std::mutex mu;
    std::vector<std::pair<std::string, int64_t>> over;
};
// One copy per block size for a graph and every share of it, each built on
// first use (SSSP sweeps and the bit-parallel batched BFS use different
// sizes, and neither rebuilds the other's).
struct ColBlockSlot {
    std::mutex mu;
    std::map<int, ColBlocks> by_shift;
    ~ColBlockSlot();
};
}  // namespace hc

struct hc_graph_s {
    int32_t n = 0;          // global vertex count
    int32_t row0 = 0;       // first row held
    int32_t rows = 0;       // rows held
    int64_t m = 0;          // edges held
    bool weighted = false;
    bool sorted = false;    // degrees non-increasing over the held rows
    bool off64 = false;
    bool shared = false;    // device arrays belong to another handle
    int32_t num_nonzero = 0;
    int64_t max_degree = 0;
    int32_t max_weight = 0;  // largest weight of an even sample of the edges
    void *d_off = nullptr;
    int32_t *d_adj = nullptr;
    int32_t *d_wt = nullptr;
    size_t bytes = 0;
    std::vector<int64_t> h_off;  // local offsets on the host (planning)
    double hub_share = -1.0;     // edge share of the top 1% rows by degree (lazy)
    std::map<std::tuple<int32_t, int32_t, hc::PlanConfig>, hc::PlanHost> plans;
    std::mutex mu;
    // Per-algorithm workspaces (allocated on first use, freed with the graph).
    void *ws_sssp = nullptr;
    void *ws_bfs = nullptr;
    void *ws_pr = nullptr;
    void *ws_multi = nullptr;
    void *ws_msbfs = nullptr;
    std::shared_ptr<hc::ColBlockSlot> cbslot = std::make_shared<hc::ColBlockSlot>();
    std::shared_ptr<hc::GraphTuning> tune = std::make_shared<hc::GraphTuning>();
};

namespace hc {
// Applies key = value to t (validated as hc_set_tuning does).
int set_knob(Tuning &t, const char *key, int64_t value);
// Makes tuning() return `t` on this thread while alive (t NULL: unchanged).
class TuningScope {
  public:
    explicit TuningScope(Tuning *t);
    ~TuningScope();
    TuningScope(const TuningScope &) = delete;
    TuningScope &operator=(const TuningScope &) = delete;

  private:
    Tuning *prev_;
    bool set_;
};
// A call on graph g: holds g->mu and, when the graph has knobs of its own,
// runs under the process-wide knobs with the graph's applied over them.
class GraphCall {
  public:
    explicit GraphCall(hc_graph_s *g);

  private:
    std::lock_guard<std::mutex> lk_;
    Tuning eff_;
    TuningScope scope_;
    static Tuning *effective(hc_graph_s *g, Tuning *eff);
};
}  // namespace hc

namespace hc {
// Plan for rows [start, end) under cfg (cached on the graph).
int get_plan(hc_graph_s *g, int32_t start, int32_t end, const PlanConfig &cfg,
             const PlanHost **out);
template <typename OffT>
inline GraphView<OffT> view(const hc_graph_s *g) {
    return GraphView<OffT>{static_cast<const OffT *>(g->d_off), g->d_adj,
                           g->d_wt, g->row0, g->n};
}
// Column blocks of 2^shift vertices for a whole-graph handle (built on
// stream s and waited for on first use); *out is NULL when the graph has
// fewer than two blocks.
int get_col_blocks(hc_graph_s *g, int shift, hipStream_t s, const ColBlocks **out);
// Block shift a pull sweep of g uses under the current tuning (0: plain
// sweep); flat: g has no hub rows (top 1% of rows hold < 5% of the edges).
int col_shift_for(const hc_graph_s *g, bool flat);
// True when the top 1% of g's rows hold under 5% of its edges (cached; the
// caller holds g->mu).
bool flat_graph(hc_graph_s *g);
void free_sssp_ws(hc_graph_s *g);
void free_bfs_ws(hc_graph_s *g);
void free_pr_ws(hc_graph_s *g);
void free_multi_ws(hc_graph_s *g);
void free_msbfs_ws(hc_graph_s *g);
// Bit-parallel BFS from up to 64 sources (whole-graph handle, caller holds
// the graph): depth[v * 64 + q] = level from sources[q], HC_INF_DIST beyond
// max_levels (<= 0: no limit) or when unreachable.
int run_msbfs_entry(hc_graph_s *g, const int32_t *sources, int num, int32_t *depth,
                    int max_levels, hc_run_stats_t *stats, hipStream_t s);
}  // namespace hc

#endif
