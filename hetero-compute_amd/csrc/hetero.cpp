// Segment-scheduled co-execution of one graph on the host cores and the GPU
// (compiled with g++ -fopenmp against the HIP runtime).
//
// The reference generates one driver per schedule from Python templates
// (src/kernels/heterogeneous/sssp.cuh:39-387, pr.cuh:39-346, bfs.cuh:19-223;
// kernelgen/sssp_hetero.py).  Here one runtime interprets any schedule:
//
//   * the graph is cut into equal-edge segments (compute_equal_edge_ranges,
//     src/graph.cuh:156-201); every schedule block owns a contiguous run of
//     segments;
//   * each GPU block keeps its rows resident in HBM as a range handle
//     (copy_subgraph_to_device, src/graph.cuh:250-286) and runs on a stream
//     of its own; all GPU blocks share one global-length state array;
//   * CPU blocks run on the host cores, on pinned host state, while the GPU
//     streams run;
//   * after every epoch the GPU-owned ranges come back to the host and, when
//     another epoch follows, the CPU-owned ranges go to the device -- one
//     PCIe exchange per epoch, as in the reference.
//
// SSSP pull epochs converge to the unique shortest-path fixpoint whatever the
// interleaving (all state only decreases), so the result is bit-exact with a
// single-device run.  PageRank runs the Jacobi form of hc_pagerank.  BFS runs
// bfs_do_cpu's direction-optimizing state machine (src/kernels/cpu/bfs.cuh:
// 48-114): top-down epochs on the host cores, bottom-up epochs split by the
// schedule between GPU and CPU blocks.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/hc_sched.h"
#include "hc_internal.h"

#define HX_HIP(expr)                                                          \
    do {                                                                      \
        hipError_t hx_e_ = (expr);                                            \
        if (hx_e_ != hipSuccess)                                              \
            return hc::fail(HC_ERR_HIP, "%s failed: %s", #expr,               \
                            hipGetErrorString(hx_e_));                        \
    } while (0)

namespace {

enum Algo { kSssp = 0, kPr = 1, kBfs = 2 };

int algo_of(int32_t kernel) {
    if (kernel <= HC_K_SSSP_GPU_BUCKETED) return kSssp;
    if (kernel <= HC_K_PR_GPU_BUCKETED) return kPr;
    return kBfs;
}

int mapping_of(int32_t kernel) {
    switch (kernel) {
        case HC_K_SSSP_GPU_ONE_TO_ONE:
        case HC_K_BFS_GPU_ONE_TO_ONE: return HC_MAP_THREAD;
        case HC_K_SSSP_GPU_WARP_RED:
        case HC_K_BFS_GPU_WARP: return HC_MAP_WAVE;
        case HC_K_SSSP_GPU_BLOCK_RED: return HC_MAP_BLOCK;
        default: return HC_MAP_BUCKETED;
    }
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Block {
    hc_schedule_block_t sb;
    int32_t start, end;          // vertex rows [start, end)
    hc_graph_t dg = nullptr;     // GPU blocks: resident rows
    hipStream_t stream = nullptr;
    int mapping = HC_MAP_BUCKETED;
};

// Host state is pinned when a GPU block exchanges it over PCIe; a schedule
// of CPU blocks only uses ordinary memory and never touches the GPU.
template <typename T>
int host_alloc(T **p, size_t count, bool pinned) {
    *p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (pinned) {
        HX_HIP(hipHostMalloc((void **)p, bytes, hipHostMallocDefault));
    } else {
        *p = (T *)std::malloc(bytes);
        if (!*p) return hc::fail(HC_ERR_OOM, "host allocation of %zu bytes failed", bytes);
    }
    return HC_OK;
}

template <typename T>
int dev_alloc(T **p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess)
        return hc::fail(e == hipErrorOutOfMemory ? HC_ERR_OOM : HC_ERR_HIP,
                        "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
                        hipGetErrorString(e));
    return HC_OK;
}

inline uint32_t saturating_relax(int32_t d, int32_t w) {
    uint32_t c = (uint32_t)d + (uint32_t)w;
    return std::min<uint32_t>(c, (uint32_t)HC_INF_DIST);
}

}  // namespace

struct hc_hetero_s {
    hc_host_graph_t g{};  // caller-owned arrays
    int algo = kSssp;
    int32_t num_segments = 0;
    std::vector<int32_t> seg;  // num_segments + 1 boundaries
    std::vector<Block> gpu, cpu;
    std::mutex mu;

    // Device state shared by every GPU block (global length).
    int32_t *d_state = nullptr;  // SSSP dist / BFS parents
    int32_t *d_cnt = nullptr;    // one counter per GPU block
    int32_t *h_state = nullptr;  // pinned
    int32_t *h_cnt = nullptr;    // pinned
    // PageRank.
    float *d_score[2] = {nullptr, nullptr}, *d_contrib[2] = {nullptr, nullptr};
    double *d_inv = nullptr, *d_resid = nullptr;
    float *h_score[2] = {nullptr, nullptr}, *h_contrib[2] = {nullptr, nullptr};
    double *h_inv = nullptr, *h_resid = nullptr;
    // BFS bitmaps (32-bit words).
    uint32_t *d_front = nullptr, *d_next = nullptr;
    uint32_t *h_front = nullptr, *h_next = nullptr, *h_gnext = nullptr;
    int32_t *h_queue = nullptr;  // top-down window, N entries
    int32_t push_kernel = HC_K_BFS_CPU_PUSH_BY_NODE;  // host top-down variant
    bool pinned = false;         // host state from hipHostMalloc

    ~hc_hetero_s() {
        for (Block &b : gpu) {
            if (b.dg) hc_graph_free(b.dg);
            if (b.stream) (void)hipStreamDestroy(b.stream);
        }
        for (void *p : {(void *)d_state, (void *)d_cnt, (void *)d_score[0],
                        (void *)d_score[1], (void *)d_contrib[0], (void *)d_contrib[1],
                        (void *)d_inv, (void *)d_resid, (void *)d_front, (void *)d_next})
            if (p) (void)hipFree(p);
        for (void *p : {(void *)h_state, (void *)h_cnt, (void *)h_score[0],
                        (void *)h_score[1], (void *)h_contrib[0], (void *)h_contrib[1],
                        (void *)h_inv, (void *)h_resid, (void *)h_front, (void *)h_next,
                        (void *)h_gnext, (void *)h_queue}) {
            if (!p) continue;
            if (pinned)
                (void)hipHostFree(p);
            else
                std::free(p);
        }
    }
};

namespace {

int sync_gpu(hc_hetero_t h) {
    for (Block &b : h->gpu) HX_HIP(hipStreamSynchronize(b.stream));
    return HC_OK;
}

hipStream_t copy_stream(hc_hetero_t h) {
    return h->gpu.empty() ? nullptr : h->gpu[0].stream;
}

// Host bottom-up epoch over [start, end): epoch_bfs_pull_cpu
// (src/kernels/cpu/bfs.cuh:305-335) with the thread's share chosen by the
// caller's team: every unvisited u takes its first neighbour (CSR order) in
// the frontier.
int64_t cpu_bfs_pull(const hc_host_graph_t &g, int32_t *parents, int32_t start,
                     int32_t end, const uint32_t *front, uint32_t *next, int tid,
                     int nthreads) {
    int64_t cnt = 0;
    for (int32_t u = start + tid; u < end; u += nthreads) {
        if (parents[u] != -1) continue;
        for (int64_t i = g.index[u]; i < g.index[u + 1]; i++) {
            int32_t v = g.neighbors[i];
            if (front[v >> 5] >> (v & 31) & 1u) {
                parents[u] = v;
                __atomic_fetch_or(&next[u >> 5], 1u << (u & 31), __ATOMIC_RELAXED);
                cnt++;
                break;
            }
        }
    }
    return cnt;
}

}  // namespace

extern "C" {

void hc_cpu_sssp_pull_epoch(const hc_host_graph_t *g, int32_t *dist,
                            int32_t start_id, int32_t end_id, int tid,
                            int nthreads, int32_t *updated) {
    int32_t local = 0;
    for (int32_t u = start_id + tid; u < end_id; u += nthreads) {
        uint32_t best = (uint32_t)__atomic_load_n(&dist[u], __ATOMIC_RELAXED);
        const uint32_t old = best;
        for (int64_t i = g->index[u]; i < g->index[u + 1]; i++) {
            int32_t dv = __atomic_load_n(&dist[g->neighbors[i]], __ATOMIC_RELAXED);
            best = std::min(best, saturating_relax(dv, g->weights[i]));
        }
        if (best < old) {
            __atomic_store_n(&dist[u], (int32_t)best, __ATOMIC_RELAXED);
            local++;
        }
    }
    if (local) __atomic_fetch_add(updated, local, __ATOMIC_RELAXED);
}

double hc_cpu_pr_pull_epoch(const hc_host_graph_t *g, float damping,
                            const float *score_in, const float *contrib,
                            const double *inv_degree, float *score_out,
                            float *contrib_out, int32_t start_id, int32_t end_id,
                            int tid, int nthreads) {
    const float base = (1.0f - damping) / (float)g->num_nodes;
    double resid = 0.0;
    for (int32_t u = start_id + tid; u < end_id; u += nthreads) {
        // Kahan sum in CSR order; value = sum - comp (the GPU sweep's KSum
        // {sum, -comp} for a single lane).  No FMA contraction: the host
        // target has none and the device code rounds each step.
        float sum = 0.0f, comp = 0.0f;
        for (int64_t i = g->index[u]; i < g->index[u + 1]; i++) {
            float y = contrib[g->neighbors[i]] - comp;
            float t = sum + y;
            comp = (t - sum) - y;
            sum = t;
        }
        float total = sum + (-comp);
        float s_new = base + damping * total;
        score_out[u] = s_new;
        contrib_out[u] = (float)((double)s_new * inv_degree[u]);
        resid += std::fabs((double)s_new - (double)score_in[u]);
    }
    return resid;
}

int hc_cpu_epoch_ms(const hc_host_graph_t *g, int32_t algorithm, int32_t start_id,
                    int32_t end_id, int32_t repeats, double *ms) {
    HC_CHECK_ARG(g && ms && g->index, "NULL argument");
    HC_CHECK_ARG(algorithm >= 0 && algorithm <= 2, "algorithm must be 0 (SSSP), 1 (BFS) or 2 (PR)");
    HC_CHECK_ARG(0 <= start_id && start_id <= end_id && end_id <= g->num_nodes,
                 "range [%d,%d) outside [0,%d)", start_id, end_id, g->num_nodes);
    HC_CHECK_ARG(algorithm != 0 || g->weights, "SSSP needs a weighted graph");
    if (repeats <= 0) repeats = 1;
    const size_t n = (size_t)g->num_nodes;
    // Work arrays in a steady state: every vertex is at distance 1 (SSSP
    // sweeps then scan whole rows without improving), every vertex unvisited
    // with an empty frontier (bottom-up scans whole rows), uniform scores.
    std::vector<int32_t> state(n, algorithm == 0 ? 1 : -1);
    std::vector<uint32_t> front((n + 31) / 32, 0), next((n + 31) / 32, 0);
    std::vector<float> score(n, 1.0f / (float)n), contrib(n, 0.0f), s_out(n), c_out(n);
    std::vector<double> inv(n, 0.0);
    for (size_t u = 0; u < n; u++) {
        const int64_t d = g->index[u + 1] - g->index[u];
        inv[u] = d > 0 ? 1.0 / (double)d : 0.0;
        contrib[u] = (float)((double)score[u] * inv[u]);
    }
    double total = 0.0;
    for (int r = 0; r < repeats; r++) {
        int32_t updated = 0;
        const double t0 = now_ms();
#pragma omp parallel
        {
            const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
            if (algorithm == 0) {
                hc_cpu_sssp_pull_epoch(g, state.data(), start_id, end_id, tid, nt, &updated);
            } else if (algorithm == 1) {
                cpu_bfs_pull(*g, state.data(), start_id, end_id, front.data(), next.data(), tid,
                             nt);
            } else {
                hc_cpu_pr_pull_epoch(g, 0.85f, score.data(), contrib.data(), inv.data(),
                                     s_out.data(), c_out.data(), start_id, end_id, tid, nt);
            }
        }
        total += now_ms() - t0;
    }
    *ms = total / repeats;
    return HC_OK;
}

int hc_hetero_create(const hc_host_graph_t *g, hc_schedule_t s, hc_hetero_t *out) {
    HC_CHECK_ARG(g && s && out, "NULL argument");
    HC_CHECK_ARG(g->index && g->num_nodes > 0, "empty graph");
    HC_CHECK_ARG(g->num_edges == 0 || g->neighbors, "graph neighbors are NULL");
    *out = nullptr;
    int32_t nb = 0, segs = 0;
    HC_TRY(hc_schedule_info(s, &nb, &segs));
    std::vector<hc_schedule_block_t> sb(nb);
    for (int32_t i = 0; i < nb; i++) HC_TRY(hc_schedule_get_block(s, i, &sb[i]));
    const int algo = algo_of(sb[0].kernel);
    // BFS push blocks (the top-down kernel the reference's push_pull_scheduler
    // assigns to a push epoch, util/scheduler/scheduler/scheduler.py:402-451)
    // pick the host-core top-down variant; they run over the whole frontier,
    // so they take no part in the segment cover of the pull blocks.  A BFS
    // schedule without pull blocks pulls on the host cores.
    int32_t push_kernel = HC_K_BFS_CPU_PUSH_BY_NODE;
    {
        std::vector<hc_schedule_block_t> pull;
        for (const hc_schedule_block_t &b : sb) {
            if (b.kernel == HC_K_BFS_CPU_PUSH_BY_NODE || b.kernel == HC_K_BFS_CPU_PUSH_BY_EDGE)
                push_kernel = b.kernel;
            else
                pull.push_back(b);
        }
        if (pull.size() != sb.size()) {
            if (pull.empty())
                pull.push_back(hc_schedule_block_t{HC_DEVICE_HOST_CPU, 0, 0, segs - 1,
                                                   HC_K_BFS_CPU_PULL, 0, 0});
            sb.swap(pull);
            nb = (int32_t)sb.size();
        }
    }
    std::vector<int> owner(segs, -1);
    for (int32_t i = 0; i < nb; i++) {
        HC_CHECK_ARG(algo_of(sb[i].kernel) == algo,
                     "block %d: kernel \"%s\" belongs to another algorithm than \"%s\"",
                     i, hc_kernel_to_string(sb[i].kernel), hc_kernel_to_string(sb[0].kernel));
        for (int32_t k = sb[i].start_segment; k <= sb[i].end_segment; k++) {
            HC_CHECK_ARG(owner[k] < 0, "segment %d is scheduled by blocks %d and %d", k,
                         owner[k], i);
            owner[k] = i;
        }
    }
    for (int32_t k = 0; k < segs; k++)
        HC_CHECK_ARG(owner[k] >= 0, "segment %d is scheduled by no block", k);
    if (algo == kSssp) HC_CHECK_ARG(g->weights, "SSSP schedules need a weighted graph");

    auto *h = new (std::nothrow) hc_hetero_s;
    if (!h) return hc::fail(HC_ERR_OOM, "out of host memory");
    std::unique_ptr<hc_hetero_s> guard(h);
    h->g = *g;
    if (algo != kSssp) h->g.weights = nullptr;
    h->algo = algo;
    h->push_kernel = push_kernel;
    h->num_segments = segs;
    h->seg.resize(size_t(segs) + 1);
    HC_TRY(hc_equal_edge_ranges(g, segs, 1, h->seg.data()));
    for (int32_t i = 0; i < nb; i++) {
        Block b;
        b.sb = sb[i];
        b.start = h->seg[sb[i].start_segment];
        b.end = h->seg[sb[i].end_segment + 1];
        int32_t dt = 0;
        HC_TRY(hc_kernel_device_type(sb[i].kernel, &dt));
        if (dt == HC_DEV_GPU) {
            b.mapping = mapping_of(sb[i].kernel);
            HC_TRY(hc_graph_upload_range(&h->g, b.start, b.end, &b.dg));
            h->gpu.push_back(b);
            HX_HIP(hipStreamCreateWithFlags(&h->gpu.back().stream, hipStreamNonBlocking));
        } else {
            h->cpu.push_back(b);
        }
    }
    const size_t n = (size_t)g->num_nodes, ng = h->gpu.size();
    const bool pin = ng > 0;
    h->pinned = pin;
    if ((algo == kSssp || algo == kBfs) && pin) {
        HC_TRY(dev_alloc(&h->d_state, n));
        HC_TRY(dev_alloc(&h->d_cnt, ng));
    }
    if (algo == kSssp || algo == kBfs) {
        HC_TRY(host_alloc(&h->h_state, n, pin));
        HC_TRY(host_alloc(&h->h_cnt, ng, pin));
    }
    if (algo == kPr) {
        for (int k = 0; k < 2; k++) {
            if (pin) HC_TRY(dev_alloc(&h->d_score[k], n));
            if (pin) HC_TRY(dev_alloc(&h->d_contrib[k], n));
            HC_TRY(host_alloc(&h->h_score[k], n, pin));
            HC_TRY(host_alloc(&h->h_contrib[k], n, pin));
        }
        if (pin) HC_TRY(dev_alloc(&h->d_inv, n));
        if (pin) HC_TRY(dev_alloc(&h->d_resid, ng));
        HC_TRY(host_alloc(&h->h_inv, n, pin));
        HC_TRY(host_alloc(&h->h_resid, ng, pin));
        for (size_t u = 0; u < n; u++) {
            int64_t d = g->index[u + 1] - g->index[u];
            h->h_inv[u] = d > 0 ? 1.0 / (double)d : 0.0;
        }
        if (pin)
            HX_HIP(hipMemcpy(h->d_inv, h->h_inv, n * sizeof(double), hipMemcpyHostToDevice));
    }
    if (algo == kBfs) {
        const size_t words = (n + 31) / 32;
        if (pin) HC_TRY(dev_alloc(&h->d_front, words));
        if (pin) HC_TRY(dev_alloc(&h->d_next, words));
        HC_TRY(host_alloc(&h->h_front, words, pin));
        HC_TRY(host_alloc(&h->h_next, words, pin));
        HC_TRY(host_alloc(&h->h_gnext, words, pin));
        HC_TRY(host_alloc(&h->h_queue, n, pin));
    }
    *out = guard.release();
    return HC_OK;
}

int hc_hetero_free(hc_hetero_t h) {
    delete h;
    return HC_OK;
}

int hc_sssp_heterogeneous(hc_hetero_t h, const int32_t *init_dist, int32_t *dist,
                          hc_run_stats_t *stats) {
    HC_CHECK_ARG(h && init_dist && dist, "NULL argument");
    HC_CHECK_ARG(h->algo == kSssp, "this schedule is not an SSSP schedule");
    std::lock_guard<std::mutex> lk(h->mu);
    const int32_t n = h->g.num_nodes;
    std::memcpy(h->h_state, init_dist, size_t(n) * 4);
    const int ng = (int)h->gpu.size();
    if (ng) HX_HIP(hipMemcpy(h->d_state, h->h_state, size_t(n) * 4, hipMemcpyHostToDevice));
    int32_t epochs = 0;
    int64_t updated = 0;
    const double t0 = now_ms();
    do {
        for (int i = 0; i < ng; i++) {
            Block &b = h->gpu[i];
            HX_HIP(hipMemsetAsync(h->d_cnt + i, 0, 4, b.stream));
            HC_TRY(hc_sssp_pull_epoch(b.dg, b.mapping, b.start, b.end, h->d_state,
                                      h->d_cnt + i, b.sb.blocks, b.sb.threads, b.stream));
            // Host cores read the GPU-owned ranges in the next epoch only
            // when the schedule has CPU blocks; otherwise the state comes
            // back once, after the loop.
            if (!h->cpu.empty())
                HX_HIP(hipMemcpyAsync(h->h_state + b.start, h->d_state + b.start,
                                      size_t(b.end - b.start) * 4, hipMemcpyDeviceToHost,
                                      b.stream));
            HX_HIP(hipMemcpyAsync(h->h_cnt + i, h->d_cnt + i, 4, hipMemcpyDeviceToHost,
                                  b.stream));
        }
        int32_t cpu_updated = 0;
        if (!h->cpu.empty()) {
#pragma omp parallel
            {
                const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
                for (const Block &b : h->cpu)
                    hc_cpu_sssp_pull_epoch(&h->g, h->h_state, b.start, b.end, tid, nt,
                                           &cpu_updated);
            }
        }
        HC_TRY(sync_gpu(h));
        updated = cpu_updated;
        for (int i = 0; i < ng; i++) updated += h->h_cnt[i];
        epochs++;
        if (updated && ng) {
            hipStream_t cs = copy_stream(h);
            for (const Block &b : h->cpu)
                HX_HIP(hipMemcpyAsync(h->d_state + b.start, h->h_state + b.start,
                                      size_t(b.end - b.start) * 4, hipMemcpyHostToDevice, cs));
            HX_HIP(hipStreamSynchronize(cs));
        }
    } while (updated != 0);
    if (ng && h->cpu.empty())
        HX_HIP(hipMemcpy(h->h_state, h->d_state, size_t(n) * 4, hipMemcpyDeviceToHost));
    const double ms = now_ms() - t0;
    std::memcpy(dist, h->h_state, size_t(n) * 4);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = ms;
        stats->epochs = epochs;
        stats->pull_epochs = epochs;
        stats->launches = (int64_t)epochs * ng;
        int64_t reached = 0;
        for (int32_t u = 0; u < n; u++) reached += dist[u] < HC_INF_DIST;
        stats->reached = reached;
        stats->alg_bytes = (double)epochs *
                           (4.0 * (n + 1) + 12.0 * (double)h->g.num_edges + 8.0 * n);
    }
    return HC_OK;
}

int hc_pr_heterogeneous(hc_hetero_t h, const hc_pr_opts_t *opts, float *scores,
                        hc_run_stats_t *stats) {
    HC_CHECK_ARG(h && scores, "NULL argument");
    HC_CHECK_ARG(h->algo == kPr, "this schedule is not a PageRank schedule");
    HC_CHECK_ARG(!opts || opts->rule == 0,
                 "host+GPU PageRank stops on the L1 residual only (rule 0)");
    const float damping = opts && opts->damping > 0 ? opts->damping : 0.85f;
    double tol = opts ? opts->tolerance : 0.0;
    if (tol == 0.0) tol = 1e-4;
    const int32_t max_iters = opts && opts->max_iters > 0 ? opts->max_iters : 20;
    HC_CHECK_ARG(damping > 0 && damping < 1, "damping must be in (0,1)");
    std::lock_guard<std::mutex> lk(h->mu);
    const int32_t n = h->g.num_nodes;
    const int ng = (int)h->gpu.size();
    const float init = 1.0f / (float)n;
    for (int32_t u = 0; u < n; u++) {
        h->h_score[0][u] = init;
        h->h_contrib[0][u] = (float)((double)init * h->h_inv[u]);
    }
    if (ng) {
        HX_HIP(hipMemcpy(h->d_score[0], h->h_score[0], size_t(n) * 4, hipMemcpyHostToDevice));
        HX_HIP(hipMemcpy(h->d_contrib[0], h->h_contrib[0], size_t(n) * 4,
                         hipMemcpyHostToDevice));
    }
    int32_t iters = 0, cur = 0;
    double residual = 0.0;
    const double t0 = now_ms();
    while (iters < max_iters) {
        const int nxt = cur ^ 1;
        for (int i = 0; i < ng; i++) {
            Block &b = h->gpu[i];
            HX_HIP(hipMemsetAsync(h->d_resid + i, 0, sizeof(double), b.stream));
            HC_TRY(hc_pr_pull_epoch(b.dg, b.start, b.end, damping, h->d_score[cur],
                                    h->d_contrib[cur], h->d_inv, h->d_score[nxt],
                                    h->d_contrib[nxt], h->d_resid + i, b.stream));
            if (!h->cpu.empty())
                HX_HIP(hipMemcpyAsync(h->h_contrib[nxt] + b.start,
                                      h->d_contrib[nxt] + b.start,
                                      size_t(b.end - b.start) * 4, hipMemcpyDeviceToHost,
                                      b.stream));
            HX_HIP(hipMemcpyAsync(h->h_resid + i, h->d_resid + i, sizeof(double),
                                  hipMemcpyDeviceToHost, b.stream));
        }
        double cpu_resid = 0.0;
        if (!h->cpu.empty()) {
            const int nthreads = omp_get_max_threads();
            std::vector<double> part(nthreads, 0.0);
#pragma omp parallel num_threads(nthreads)
            {
                const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
                double r = 0.0;
                for (const Block &b : h->cpu)
                    r += hc_cpu_pr_pull_epoch(&h->g, damping, h->h_score[cur],
                                              h->h_contrib[cur], h->h_inv, h->h_score[nxt],
                                              h->h_contrib[nxt], b.start, b.end, tid, nt);
                part[tid] = r;
            }
            for (double r : part) cpu_resid += r;  // fixed order
        }
        HC_TRY(sync_gpu(h));
        residual = cpu_resid;
        for (int i = 0; i < ng; i++) residual += h->h_resid[i];
        iters++;
        cur = nxt;
        if (tol > 0 && residual < tol) break;
        if (iters < max_iters && ng && !h->cpu.empty()) {
            hipStream_t cs = copy_stream(h);
            for (const Block &b : h->cpu)
                HX_HIP(hipMemcpyAsync(h->d_contrib[cur] + b.start, h->h_contrib[cur] + b.start,
                                      size_t(b.end - b.start) * 4, hipMemcpyHostToDevice, cs));
            HX_HIP(hipStreamSynchronize(cs));
        }
    }
    const double ms = now_ms() - t0;
    // Scores: GPU ranges from the device, CPU ranges from the host.
    for (const Block &b : h->gpu)
        HX_HIP(hipMemcpy(scores + b.start, h->d_score[cur] + b.start,
                         size_t(b.end - b.start) * 4, hipMemcpyDeviceToHost));
    for (const Block &b : h->cpu)
        std::memcpy(scores + b.start, h->h_score[cur] + b.start, size_t(b.end - b.start) * 4);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = ms;
        stats->epochs = iters;
        stats->pull_epochs = iters;
        stats->launches = (int64_t)iters * ng;
        stats->reached = n;
        stats->residual = residual;
        stats->alg_bytes = (double)iters *
                           (4.0 * (n + 1) + 8.0 * (double)h->g.num_edges + 28.0 * n);
    }
    return HC_OK;
}

int hc_bfs_heterogeneous(hc_hetero_t h, int32_t source, int32_t alpha, int32_t beta,
                         int32_t *parents, int32_t *depth, hc_run_stats_t *stats) {
    HC_CHECK_ARG(h && parents, "NULL argument");
    HC_CHECK_ARG(h->algo == kBfs, "this schedule is not a BFS schedule");
    const int32_t n = h->g.num_nodes;
    HC_CHECK_ARG(source >= 0 && source < n, "source %d outside [0,%d)", source, n);
    if (alpha <= 0) alpha = 15;
    if (beta <= 0) beta = 18;
    std::lock_guard<std::mutex> lk(h->mu);
    const hc_host_graph_t &g = h->g;
    const int ng = (int)h->gpu.size();
    const size_t words = (size_t(n) + 31) / 32;
    int32_t *par = h->h_state;
    int32_t *q = h->h_queue;
    for (int32_t u = 0; u < n; u++) par[u] = -1;
    if (depth)
        for (int32_t u = 0; u < n; u++) depth[u] = -1;
    par[source] = source;
    if (depth) depth[source] = 0;
    std::memset(h->h_front, 0, words * 4);
    std::memset(h->h_next, 0, words * 4);
    int64_t w_start = 0, w_end = 1;
    q[0] = source;
    int64_t edges_to_check = g.num_edges;
    int64_t num_edges = g.index[source + 1] - g.index[source];
    int32_t level = 0, epochs = 0, push_epochs = 0, pull_epochs = 0;
    int64_t launches = 0;
    const double t0 = now_ms();
    while (w_start != w_end) {
        if (num_edges > edges_to_check / alpha) {
            // Bottom-up phase (bfs_do_cpu's pull loop, :72-89).
            int64_t num_nodes = w_end - w_start, prev_nodes;
            for (int64_t i = w_start; i < w_end; i++)
                h->h_front[q[i] >> 5] |= 1u << (q[i] & 31);
            if (ng) {
                hipStream_t cs = copy_stream(h);
                for (const Block &b : h->gpu)
                    HX_HIP(hipMemcpyAsync(h->d_state + b.start, par + b.start,
                                          size_t(b.end - b.start) * 4, hipMemcpyHostToDevice,
                                          cs));
                HX_HIP(hipMemcpyAsync(h->d_front, h->h_front, words * 4,
                                      hipMemcpyHostToDevice, cs));
                HX_HIP(hipMemsetAsync(h->d_next, 0, words * 4, cs));
                HX_HIP(hipStreamSynchronize(cs));
            }
            do {
                prev_nodes = num_nodes;
                level++;
                for (int i = 0; i < ng; i++) {
                    Block &b = h->gpu[i];
                    HX_HIP(hipMemsetAsync(h->d_cnt + i, 0, 4, b.stream));
                    HC_TRY(hc_bfs_pull_epoch(b.dg, b.mapping, b.start, b.end, h->d_state,
                                             h->d_front, h->d_next, h->d_cnt + i,
                                             b.sb.blocks, b.sb.threads, b.stream));
                    HX_HIP(hipMemcpyAsync(h->h_cnt + i, h->d_cnt + i, 4,
                                          hipMemcpyDeviceToHost, b.stream));
                }
                launches += ng;
                int64_t cnt = 0;
                if (!h->cpu.empty()) {
#pragma omp parallel reduction(+ : cnt)
                    {
                        const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
                        for (const Block &b : h->cpu)
                            cnt += cpu_bfs_pull(g, par, b.start, b.end, h->h_front,
                                                h->h_next, tid, nt);
                    }
                }
                HC_TRY(sync_gpu(h));
                for (int i = 0; i < ng; i++) cnt += h->h_cnt[i];
                num_nodes = cnt;
                if (ng) {
                    // Merge the GPU blocks' bits into the host's next frontier.
                    hipStream_t cs = copy_stream(h);
                    HX_HIP(hipMemcpyAsync(h->h_gnext, h->d_next, words * 4,
                                          hipMemcpyDeviceToHost, cs));
                    HX_HIP(hipStreamSynchronize(cs));
                    for (size_t w = 0; w < words; w++) h->h_next[w] |= h->h_gnext[w];
                }
                if (depth)
                    for (size_t w = 0; w < words; w++)
                        for (uint32_t bits = h->h_next[w]; bits; bits &= bits - 1)
                            depth[w * 32 + __builtin_ctz(bits)] = level;
                std::swap(h->h_front, h->h_next);
                std::memset(h->h_next, 0, words * 4);
                epochs++;
                pull_epochs++;
                const bool again = num_nodes > prev_nodes || num_nodes > n / beta;
                if (again && ng) {
                    hipStream_t cs = copy_stream(h);
                    HX_HIP(hipMemcpyAsync(h->d_front, h->h_front, words * 4,
                                          hipMemcpyHostToDevice, cs));
                    HX_HIP(hipMemsetAsync(h->d_next, 0, words * 4, cs));
                    HX_HIP(hipStreamSynchronize(cs));
                }
                if (!again) break;
            } while (true);
            // Parents of GPU-owned rows back to the host for the top-down
            // epochs and the result.
            if (ng) {
                hipStream_t cs = copy_stream(h);
                for (const Block &b : h->gpu)
                    HX_HIP(hipMemcpyAsync(par + b.start, h->d_state + b.start,
                                          size_t(b.end - b.start) * 4, hipMemcpyDeviceToHost,
                                          cs));
                HX_HIP(hipStreamSynchronize(cs));
            }
            num_edges = 1;
            // conv_bitmap_to_window (:361-379): ascending vertex order.
            w_start = w_end = 0;
            for (size_t w = 0; w < words; w++)
                for (uint32_t bits = h->h_front[w]; bits; bits &= bits - 1)
                    q[w_end++] = (int32_t)(w * 32 + __builtin_ctz(bits));
            std::memset(h->h_front, 0, words * 4);
        } else {
            // Top-down epoch on the host cores: frontier vertices split
            // between threads (epoch_bfs_push_cpu_by_node, :212-244), or
            // each vertex's edges split between them when the schedule names
            // the by-edge kernel (epoch_bfs_push_cpu_by_edge, :255-288).
            edges_to_check -= num_edges;
            level++;
            int64_t ne = 0;
            std::atomic<int64_t> tail{w_end};
            const int64_t fs = w_start, fe = w_end;
            const bool by_edge = h->push_kernel == HC_K_BFS_CPU_PUSH_BY_EDGE;
#pragma omp parallel reduction(+ : ne)
            {
                constexpr int kLocal = 4096;
                int32_t loc[kLocal];
                int lc = 0;
                auto visit = [&](int32_t u, int64_t i) {
                    const int32_t v = g.neighbors[i];
                    int32_t expect = -1;
                    if (__atomic_load_n(&par[v], __ATOMIC_RELAXED) == -1 &&
                        __atomic_compare_exchange_n(&par[v], &expect, u, false,
                                                    __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                        if (lc == kLocal) {
                            int64_t at = tail.fetch_add(lc);
                            std::memcpy(q + at, loc, sizeof(int32_t) * lc);
                            lc = 0;
                        }
                        loc[lc++] = v;
                        if (depth) depth[v] = level;
                        ne += g.index[v + 1] - g.index[v];
                    }
                };
                if (by_edge) {
                    for (int64_t f = fs; f < fe; f++) {
                        const int32_t u = q[f];
                        const int64_t b = g.index[u], e = g.index[u + 1];
#pragma omp for schedule(static) nowait
                        for (int64_t i = b; i < e; i++) visit(u, i);
                    }
                } else {
#pragma omp for schedule(dynamic, 64) nowait
                    for (int64_t f = fs; f < fe; f++) {
                        const int32_t u = q[f];
                        for (int64_t i = g.index[u]; i < g.index[u + 1]; i++) visit(u, i);
                    }
                }
                int64_t at = tail.fetch_add(lc);
                std::memcpy(q + at, loc, sizeof(int32_t) * lc);
            }
            num_edges = ne;
            w_start = fe;
            w_end = tail.load();
            epochs++;
            push_epochs++;
        }
    }
    const double ms = now_ms() - t0;
    std::memcpy(parents, par, size_t(n) * 4);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = ms;
        stats->epochs = epochs;
        stats->push_epochs = push_epochs;
        stats->pull_epochs = pull_epochs;
        stats->launches = launches;
        int64_t reached = 0;
        for (int32_t u = 0; u < n; u++) reached += par[u] != -1;
        stats->reached = reached;
    }
    return HC_OK;
}

}  // extern "C"
