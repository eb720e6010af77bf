// C++ drop-in for the reference's kernel registries and drivers:
//   GPU  src/kernels/gpu/sssp.cuh:42-100,288-356; gpu/bfs.cuh:21-95,295-365;
//        gpu/pr.cuh:42-118,325-393
//   CPU  src/kernels/cpu/sssp.cuh:33-65,155-251; cpu/bfs.cuh:48-114,404-531;
//        cpu/pr.cuh:33-65,157-256
//   heterogeneous  src/kernels/heterogeneous/{sssp,pr,bfs}.cuh
//   epoch-kernel types  src/kernels/kernel_types.cuh:32-63
// over the C ABI of libhcamd.so.  Header-only (OpenMP host code); link with
// -lhcamd -fopenmp.
//
// Same driver names, argument order, return value (elapsed ms of the epoch
// loop) and ownership (results are new[]-allocated for the caller to
// delete[]), and the reference's print-and-exit on failure.  Where the
// reference passes a __global__ function pointer, a registry id is passed:
// get_kernel(id) returns the id itself, so `get_kernel(SSSPGPU::warp_red)`
// still reads the same at the call site.  Each call uploads the graph, as
// the reference drivers do; DeviceGraph keeps one resident instead.
#ifndef HC_CPP_KERNELS_H
#define HC_CPP_KERNELS_H

#include <omp.h>

#include <chrono>
#include <cmath>
#include <functional>
#include <memory>
#include <ostream>
#include <queue>
#include <string>
#include <utility>
#include <vector>

#include "../hc_batch.h"
#include "../hc_sched.h"
#include "bitmap.h"
#include "graph.h"
#include "window.h"

#ifndef UNUSED
#define UNUSED __attribute__((unused))
#endif

// ---------------------------------------------------------------------------
// Registries (names as the reference's to_repr / to_string)
// ---------------------------------------------------------------------------

enum class SSSPCPU { one_to_one, undefined };
enum class SSSPGPU { one_to_one, warp_red, block_red, bucketed, undefined };
enum class PRCPU { one_to_one, undefined };
enum class PRGPU { one_to_one, warp_red, block_red, bucketed, undefined };
enum class BFSCPUPush { by_node, by_edge, undefined };
enum class BFSCPUPull { pull, undefined };
// sync_warp: epoch_bfs_sync_pull_gpu_warp (src/kernels/gpu/bfs.cuh:234-288),
// which the reference leaves out of its registry; it runs the wavefront
// mapping, whose vertices stop at the first round of 64 edges that holds a
// frontier neighbour (the sync_iters = 1 form).
enum class BFSGPU { one_to_one, warp, bucketed, sync_warp, undefined };

inline std::vector<SSSPCPU> get_kernels(UNUSED SSSPCPU u) { return {SSSPCPU::one_to_one}; }
inline std::vector<SSSPGPU> get_kernels(UNUSED SSSPGPU u) {
    return {SSSPGPU::one_to_one, SSSPGPU::warp_red, SSSPGPU::block_red, SSSPGPU::bucketed};
}
inline std::vector<PRCPU> get_kernels(UNUSED PRCPU u) { return {PRCPU::one_to_one}; }
inline std::vector<PRGPU> get_kernels(UNUSED PRGPU u) {
    return {PRGPU::one_to_one, PRGPU::warp_red, PRGPU::block_red, PRGPU::bucketed};
}
inline std::vector<BFSCPUPush> get_kernels(UNUSED BFSCPUPush u) {
    return {BFSCPUPush::by_node, BFSCPUPush::by_edge};
}
inline std::vector<BFSCPUPull> get_kernels(UNUSED BFSCPUPull u) { return {BFSCPUPull::pull}; }
inline std::vector<BFSGPU> get_kernels(UNUSED BFSGPU u) {
    return {BFSGPU::one_to_one, BFSGPU::warp, BFSGPU::bucketed};
}

inline std::string to_repr(SSSPCPU k) { return k == SSSPCPU::one_to_one ? "sssp_cpu_onetoone" : ""; }
inline std::string to_repr(SSSPGPU k) {
    switch (k) {
        case SSSPGPU::one_to_one: return "sssp_gpu_onetoone";
        case SSSPGPU::warp_red: return "sssp_gpu_warp_red";
        case SSSPGPU::block_red: return "sssp_gpu_block_red";
        case SSSPGPU::bucketed: return "sssp_gpu_bucketed";
        default: return "";
    }
}
inline std::string to_repr(PRCPU k) { return k == PRCPU::one_to_one ? "pr_cpu_onetoone" : ""; }
inline std::string to_repr(PRGPU k) {
    switch (k) {
        case PRGPU::one_to_one: return "pr_gpu_onetoone";
        case PRGPU::warp_red: return "pr_gpu_warp_red";
        case PRGPU::block_red: return "pr_gpu_block_red";
        case PRGPU::bucketed: return "pr_gpu_bucketed";
        default: return "";
    }
}
inline std::string to_repr(BFSCPUPush k) {
    switch (k) {
        case BFSCPUPush::by_node: return "bfs_cpu_push_by_node";
        case BFSCPUPush::by_edge: return "bfs_cpu_push_by_edge";
        default: return "";
    }
}
inline std::string to_repr(BFSCPUPull k) { return k == BFSCPUPull::pull ? "bfs_cpu_pull" : ""; }
inline std::string to_repr(BFSGPU k) {
    switch (k) {
        case BFSGPU::one_to_one: return "bfs_gpu_onetoone";
        case BFSGPU::warp: return "bfs_gpu_warp";
        case BFSGPU::bucketed: return "bfs_gpu_bucketed";
        case BFSGPU::sync_warp: return "bfs_gpu_sync_warp";
        default: return "";
    }
}

// Human-readable names: the strings .skd schedules carry (hc_sched.h).
inline int32_t kernel_id_of(SSSPCPU k) { return k == SSSPCPU::one_to_one ? HC_K_SSSP_CPU_ONE_TO_ONE : -1; }
inline int32_t kernel_id_of(SSSPGPU k) {
    return k == SSSPGPU::undefined ? -1 : HC_K_SSSP_GPU_ONE_TO_ONE + (int32_t)k;
}
inline int32_t kernel_id_of(PRCPU k) { return k == PRCPU::one_to_one ? HC_K_PR_CPU_ONE_TO_ONE : -1; }
inline int32_t kernel_id_of(PRGPU k) {
    return k == PRGPU::undefined ? -1 : HC_K_PR_GPU_ONE_TO_ONE + (int32_t)k;
}
inline int32_t kernel_id_of(BFSCPUPull k) { return k == BFSCPUPull::pull ? HC_K_BFS_CPU_PULL : -1; }
inline int32_t kernel_id_of(BFSCPUPush k) {
    switch (k) {
        case BFSCPUPush::by_node: return HC_K_BFS_CPU_PUSH_BY_NODE;
        case BFSCPUPush::by_edge: return HC_K_BFS_CPU_PUSH_BY_EDGE;
        default: return -1;
    }
}
inline int32_t kernel_id_of(BFSGPU k) {
    switch (k) {
        case BFSGPU::one_to_one: return HC_K_BFS_GPU_ONE_TO_ONE;
        case BFSGPU::warp:
        case BFSGPU::sync_warp: return HC_K_BFS_GPU_WARP;
        case BFSGPU::bucketed: return HC_K_BFS_GPU_BUCKETED;
        default: return -1;
    }
}
inline std::string to_string(SSSPCPU k) {
    return k == SSSPCPU::undefined ? "undefined SSSP CPU kernel" : hc_kernel_to_string(kernel_id_of(k));
}
inline std::string to_string(SSSPGPU k) {
    return k == SSSPGPU::undefined ? "undefined SSSP GPU kernel" : hc_kernel_to_string(kernel_id_of(k));
}
inline std::string to_string(PRCPU k) {
    return k == PRCPU::undefined ? "undefined PR CPU kernel" : hc_kernel_to_string(kernel_id_of(k));
}
inline std::string to_string(PRGPU k) {
    return k == PRGPU::undefined ? "undefined PR GPU kernel" : hc_kernel_to_string(kernel_id_of(k));
}
// The reference swaps these two labels (src/kernels/cpu/bfs.cuh:470-471);
// here each name matches its kernel.
inline std::string to_string(BFSCPUPush k) {
    switch (k) {
        case BFSCPUPush::by_node: return "BFS CPU push by node";
        case BFSCPUPush::by_edge: return "BFS CPU push by edge";
        default: return "undefined BFS CPU push kernel";
    }
}
inline std::string to_string(BFSCPUPull k) {
    return k == BFSCPUPull::undefined ? "undefined BFS CPU pull kernel" : hc_kernel_to_string(kernel_id_of(k));
}
inline std::string to_string(BFSGPU k) {
    if (k == BFSGPU::sync_warp) return "BFS GPU sync warp";
    return k == BFSGPU::undefined ? "undefined BFS GPU kernel" : hc_kernel_to_string(kernel_id_of(k));
}

template <typename K>
inline K get_kernel(K k) { return k; }

#define HC_DEFINE_KERNEL_OSTREAM(T) \
    inline std::ostream &operator<<(std::ostream &os, T k) { return os << to_string(k); }
HC_DEFINE_KERNEL_OSTREAM(SSSPCPU)
HC_DEFINE_KERNEL_OSTREAM(SSSPGPU)
HC_DEFINE_KERNEL_OSTREAM(PRCPU)
HC_DEFINE_KERNEL_OSTREAM(PRGPU)
HC_DEFINE_KERNEL_OSTREAM(BFSCPUPush)
HC_DEFINE_KERNEL_OSTREAM(BFSCPUPull)
HC_DEFINE_KERNEL_OSTREAM(BFSGPU)
#undef HC_DEFINE_KERNEL_OSTREAM

// Epoch-kernel types (src/kernels/kernel_types.cuh:32-68): CPU epochs keep
// the reference signatures (PageRank scores are float here, where the
// reference's are its float weight_t); GPU epochs are registry ids.
typedef void (*sssp_cpu_epoch_func)(const CSRWGraph &, weight_t *, const nid_t, const nid_t,
                                    const int, const int, nid_t &);
typedef void (*pr_cpu_epoch_func)(const CSRWGraph &, float *, const nid_t, const nid_t,
                                  const int, const int, nid_t &);
typedef void (*bfs_cpu_push_epoch_func)(const CSRUWGraph &, nid_t *const,
                                        SlidingWindow<nid_t> &, nid_t &);
typedef void (*bfs_cpu_pull_epoch_func)(const CSRUWGraph &, nid_t *const, const nid_t,
                                        const nid_t, const Bitmap::Bitmap *const,
                                        Bitmap::Bitmap *const, nid_t &);
typedef double(bfs_cpu_kernel)(const CSRUWGraph &g, const nid_t, nid_t **const);
typedef SSSPGPU sssp_gpu_epoch_func;
typedef PRGPU pr_gpu_epoch_func;
typedef BFSGPU bfs_gpu_epoch_func;

// The reference's GPU epoch kernels by name (src/kernels/gpu/sssp.cuh:120-
// 281, gpu/pr.cuh:140-318, gpu/bfs.cuh:116-288), as the registry ids the
// drivers take, so call sites such as
// sssp_pull_gpu(g, epoch_sssp_pull_gpu_warp_red, init, &dist) read the same.
constexpr SSSPGPU epoch_sssp_pull_gpu_one_to_one = SSSPGPU::one_to_one;
constexpr SSSPGPU epoch_sssp_pull_gpu_warp_red = SSSPGPU::warp_red;
constexpr SSSPGPU epoch_sssp_pull_gpu_block_red = SSSPGPU::block_red;
constexpr PRGPU epoch_pr_pull_gpu_one_to_one = PRGPU::one_to_one;
constexpr PRGPU epoch_pr_pull_gpu_warp_red = PRGPU::warp_red;
constexpr PRGPU epoch_pr_pull_gpu_block_red = PRGPU::block_red;
constexpr BFSGPU epoch_bfs_pull_gpu_one_to_one = BFSGPU::one_to_one;
constexpr BFSGPU epoch_bfs_pull_gpu_warp = BFSGPU::warp;
constexpr BFSGPU epoch_bfs_sync_pull_gpu_warp = BFSGPU::sync_warp;

namespace hc {

inline int mapping_of(SSSPGPU k) {
    switch (k) {
        case SSSPGPU::one_to_one: return HC_MAP_THREAD;
        case SSSPGPU::warp_red: return HC_MAP_WAVE;
        case SSSPGPU::block_red: return HC_MAP_BLOCK;
        default: return HC_MAP_BUCKETED;
    }
}
inline int mapping_of(PRGPU k) {
    switch (k) {
        case PRGPU::one_to_one: return HC_MAP_THREAD;
        case PRGPU::warp_red: return HC_MAP_WAVE;
        case PRGPU::block_red: return HC_MAP_BLOCK;
        default: return HC_MAP_BUCKETED;
    }
}
inline int mapping_of(BFSGPU k) {
    switch (k) {
        case BFSGPU::one_to_one: return HC_MAP_THREAD;
        case BFSGPU::warp:
        case BFSGPU::sync_warp: return HC_MAP_WAVE;
        default: return HC_MAP_BUCKETED;
    }
}

// A graph resident in HBM for as long as the object lives.
class DeviceGraph {
public:
    template <typename NeighborT>
    explicit DeviceGraph(const GenericCSRGraph<NeighborT> &g) {
        check_or_exit(hc_init(0), "hc_init");
        HostView<NeighborT> v(g);
        check_or_exit(hc_graph_upload(&v.c, &h), "hc_graph_upload");
    }
    ~DeviceGraph() {
        if (h) hc_graph_free(h);
    }
    DeviceGraph(const DeviceGraph &) = delete;
    DeviceGraph &operator=(const DeviceGraph &) = delete;
    hc_graph_t get() const { return h; }

private:
    hc_graph_t h = nullptr;
};

template <typename T>
struct DeviceBuffer {
    T *p = nullptr;
    size_t n = 0;
    explicit DeviceBuffer(size_t count) : n(count) {
        void *q = nullptr;
        check_or_exit(hc_dmalloc(&q, sizeof(T) * (count ? count : 1)), "hc_dmalloc");
        p = static_cast<T *>(q);
    }
    ~DeviceBuffer() { hc_dfree(p); }
    void from(const T *src) { check_or_exit(hc_memcpy(p, src, sizeof(T) * n, 1, nullptr), "hc_memcpy"); }
    void to(T *dst) const { check_or_exit(hc_memcpy(dst, p, sizeof(T) * n, 2, nullptr), "hc_memcpy"); }
};

inline double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace hc

// ---------------------------------------------------------------------------
// GPU drivers
// ---------------------------------------------------------------------------

// sssp_pull_gpu (src/kernels/gpu/sssp.cuh:42-100): pull epochs of the given
// mapping from init_dist until no vertex changes.
inline double sssp_pull_gpu(const CSRWGraph &g, sssp_gpu_epoch_func epoch_kernel,
                            const weight_t *init_dist, weight_t **const ret_dist,
                            int block_count = 64, int thread_count = 1024) {
    hc::DeviceGraph dg(g);
    hc::DeviceBuffer<weight_t> init(g.num_nodes), dist(g.num_nodes);
    init.from(init_dist);
    hc_run_stats_t st{};
    hc::check_or_exit(hc_sssp_pull(dg.get(), hc::mapping_of(epoch_kernel), init.p, dist.p,
                                   block_count, thread_count, &st, nullptr),
                      "sssp_pull_gpu");
    *ret_dist = new weight_t[g.num_nodes];
    dist.to(*ret_dist);
    return st.ms;
}

// Direction-optimizing SSSP from one source (push/pull epochs); the
// library's fastest single-source path.
inline double sssp_gpu(const CSRWGraph &g, nid_t source_id, weight_t **const ret_dist) {
    hc::DeviceGraph dg(g);
    hc::DeviceBuffer<weight_t> dist(g.num_nodes);
    hc_run_stats_t st{};
    hc::check_or_exit(hc_sssp(dg.get(), source_id, dist.p, nullptr, &st, nullptr), "sssp_gpu");
    *ret_dist = new weight_t[g.num_nodes];
    dist.to(*ret_dist);
    return st.ms;
}

// bfs_gpu (src/kernels/gpu/bfs.cuh:21-95): bottom-up epochs only, with the
// given mapping, from the source until no vertex joins; returns parents
// (new[]).  Each vertex takes its first frontier neighbour in CSR order.
inline double bfs_gpu(const CSRUWGraph &g, const nid_t source_id, nid_t **const ret_parents,
                      bfs_gpu_epoch_func epoch_kernel, int block_count = 64,
                      int thread_count = 1024) {
    hc::DeviceGraph dg(g);
    const size_t words = ((size_t)g.num_nodes + 31) / 32;
    hc::DeviceBuffer<nid_t> parents(g.num_nodes), count(1);
    hc::DeviceBuffer<uint32_t> front(words), next(words);
    std::vector<nid_t> p(g.num_nodes, INVALID_NODE);
    p[source_id] = source_id;
    parents.from(p.data());
    std::vector<uint32_t> f(words, 0);
    front.from(f.data());
    next.from(f.data());
    f[source_id >> 5] = 1u << (source_id & 31);
    front.from(f.data());
    const int mapping = hc::mapping_of(epoch_kernel);
    nid_t n = 1;
    const double t0 = hc::now_ms();
    while (n != 0) {
        hc::check_or_exit(hc_memset_d32(count.p, 0, 1, nullptr), "hc_memset_d32");
        hc::check_or_exit(hc_bfs_pull_epoch(dg.get(), mapping, 0, g.num_nodes, parents.p, front.p,
                                            next.p, count.p, block_count, thread_count, nullptr),
                          "bfs_gpu");
        count.to(&n);
        std::swap(front.p, next.p);
        hc::check_or_exit(hc_memset_d32(next.p, 0, words, nullptr), "hc_memset_d32");
    }
    const double ms = hc::now_ms() - t0;
    *ret_parents = new nid_t[g.num_nodes];
    parents.to(*ret_parents);
    return ms;
}

// Direction-optimizing BFS on the GPU (bfs_do_cpu's switching rule, top-down
// and bottom-up epochs in one device-driven loop): the library's fast path.
template <int alpha = 15, int beta = 18>
inline double bfs_do_gpu(const CSRUWGraph &g, const nid_t source_id, nid_t **const ret_parents) {
    hc::DeviceGraph dg(g);
    hc::DeviceBuffer<nid_t> depth(g.num_nodes), parents(g.num_nodes);
    hc_bfs_opts_t o{alpha, beta, 0};
    hc_run_stats_t st{};
    hc::check_or_exit(hc_bfs(dg.get(), source_id, depth.p, parents.p, &o, &st, nullptr),
                      "bfs_do_gpu");
    *ret_parents = new nid_t[g.num_nodes];
    parents.to(*ret_parents);
    return st.ms;
}

// pr_pull_gpu (src/kernels/gpu/pr.cuh:42-118): the given epoch kernel
// (a thread / wavefront / workgroup per vertex, or the degree-bucketed
// sweep) on block_count x thread_count workgroups (thread_count a multiple of
// 64 except for one_to_one, as the reference asks multiples of its warp),
// starting from init_score (host; NULL: 1/N), until an iteration in which no
// vertex moved by more than 5e-8, at most 201 iterations: the reference's
// loop and stopping rule (:83-96).  The iteration is Jacobi with FP32 Kahan
// sums and FP64 1/deg, where the reference's kernels update in place.
template <typename GraphT>
inline double pr_pull_gpu(const GraphT &g, pr_gpu_epoch_func epoch_kernel, float *init_score,
                          float **const ret_score, int block_count = 64,
                          int thread_count = 1024) {
    hc::DeviceGraph dg(g);
    hc::DeviceBuffer<float> score(g.num_nodes);
    std::unique_ptr<hc::DeviceBuffer<float>> init;
    if (init_score) {
        init.reset(new hc::DeviceBuffer<float>(g.num_nodes));
        init->from(init_score);
    }
    hc_pr_opts_t o{0.85f, 0.0, 201, 1};
    hc_run_stats_t st{};
    hc::check_or_exit(hc_pr_pull(dg.get(), hc::mapping_of(epoch_kernel), init ? init->p : nullptr,
                                 score.p, block_count, thread_count, &o, &st, nullptr),
                      "pr_pull_gpu");
    *ret_score = new float[g.num_nodes];
    score.to(*ret_score);
    return st.ms;
}

// Direction-optimizing-free PageRank on the GPU at the library's defaults
// (FP64 L1 residual < 1e-4, <= 20 iterations, fused degree-bucketed sweep).
template <typename GraphT>
inline double pr_gpu(const GraphT &g, float **const ret_score) {
    hc::DeviceGraph dg(g);
    hc::DeviceBuffer<float> score(g.num_nodes);
    hc_pr_opts_t o{0.85f, 1e-4, 20, 0};
    hc_run_stats_t st{};
    hc::check_or_exit(hc_pagerank(dg.get(), score.p, &o, &st, nullptr), "pr_gpu");
    *ret_score = new float[g.num_nodes];
    score.to(*ret_score);
    return st.ms;
}

// ---------------------------------------------------------------------------
// CPU epochs and drivers (host cores; OpenMP)
// ---------------------------------------------------------------------------

// epoch_sssp_pull_cpu_one_to_one (src/kernels/cpu/sssp.cuh:155-182): called
// by every thread of a parallel region.
inline void epoch_sssp_pull_cpu_one_to_one(const CSRWGraph &g, weight_t *dist,
                                           const nid_t start_id, const nid_t end_id,
                                           const int tid, const int num_threads,
                                           nid_t &updated) {
    nid_t local = 0;
    for (nid_t u = start_id + tid; u < end_id; u += num_threads) {
        weight_t best = __atomic_load_n(&dist[u], __ATOMIC_RELAXED);
        const weight_t old = best;
        for (wnode_t nei : g.get_neighbors(u)) {
            const uint32_t c = (uint32_t)__atomic_load_n(&dist[nei.v], __ATOMIC_RELAXED) +
                               (uint32_t)nei.w;
            const weight_t cand = c < (uint32_t)INF_WEIGHT ? (weight_t)c : INF_WEIGHT;
            if (cand < best) best = cand;
        }
        if (best < old) {
            __atomic_store_n(&dist[u], best, __ATOMIC_RELAXED);
            local++;
        }
    }
    if (local) __atomic_fetch_add(&updated, local, __ATOMIC_RELAXED);
}

// get_kernel for the CPU registry returns the epoch function, as the
// reference's does (src/kernels/cpu/sssp.cuh:245-251).
inline sssp_cpu_epoch_func get_kernel(SSSPCPU k) {
    return k == SSSPCPU::one_to_one ? epoch_sssp_pull_cpu_one_to_one : nullptr;
}

// sssp_pull_cpu (src/kernels/cpu/sssp.cuh:33-65).
inline double sssp_pull_cpu(const CSRWGraph &g, sssp_cpu_epoch_func epoch_kernel,
                            const weight_t *init_dist, weight_t **const ret_dist) {
    weight_t *dist = new weight_t[g.num_nodes];
#pragma omp parallel for
    for (nid_t i = 0; i < g.num_nodes; i++) dist[i] = init_dist[i];
    nid_t updated = 1;
    const double t0 = hc::now_ms();
    while (updated != 0) {
        updated = 0;
#pragma omp parallel
        epoch_kernel(g, dist, 0, g.num_nodes, omp_get_thread_num(), omp_get_num_threads(),
                     updated);
    }
    const double ms = hc::now_ms() - t0;
    *ret_dist = dist;
    return ms;
}

namespace hc {

// A one-block CPU schedule over the whole graph for the co-execution runtime.
template <typename NeighborT>
inline double run_cpu_schedule(const GenericCSRGraph<NeighborT> &g, int32_t kernel,
                               hc_run_stats_t *st, void (*run)(hc_hetero_t, hc_run_stats_t *, void *),
                               void *arg) {
    hc_schedule_block_t b{HC_DEVICE_HOST_CPU, 0, 0, 0, kernel, 0, 0};
    hc_schedule_t s = nullptr;
    check_or_exit(hc_schedule_create(&b, 1, 1, &s), "hc_schedule_create");
    HostView<NeighborT> v(g);
    hc_hetero_t h = nullptr;
    check_or_exit(hc_hetero_create(&v.c, s, &h), "hc_hetero_create");
    run(h, st, arg);
    hc_hetero_free(h);
    hc_schedule_free(s);
    return st->ms;
}

}  // namespace hc

// bfs_do_cpu<alpha, beta> (src/kernels/cpu/bfs.cuh:48-114) on the host cores.
template <int alpha = 15, int beta = 18>
inline double bfs_do_cpu(const CSRUWGraph &g, const nid_t source_id, nid_t **const ret_parents) {
    nid_t *parents = new nid_t[g.num_nodes];
    struct A {
        nid_t src;
        nid_t *par;
    } a{source_id, parents};
    hc_run_stats_t st{};
    hc::run_cpu_schedule(g, HC_K_BFS_CPU_PULL, &st,
                         [](hc_hetero_t h, hc_run_stats_t *s, void *p) {
                             A *x = static_cast<A *>(p);
                             hc::check_or_exit(hc_bfs_heterogeneous(h, x->src, alpha, beta,
                                                                    x->par, nullptr, s),
                                               "bfs_do_cpu");
                         },
                         &a);
    *ret_parents = parents;
    return st.ms;
}

// ---- serial SSSP oracles (src/kernels/cpu/sssp.cuh:74-137) ----

// Bellman-Ford sweeps in vertex order, in place, until one changes nothing.
inline void sssp_pull_cpu_serial(const CSRWGraph &g, nid_t source_id, weight_t **ret_dist) {
    weight_t *dist = new weight_t[g.num_nodes];
    std::fill(dist, dist + g.num_nodes, INF_WEIGHT);
    dist[source_id] = 0;
    bool changed = true;
    while (changed) {
        changed = false;
        for (nid_t u = 0; u < g.num_nodes; u++)
            for (wnode_t nei : g.get_neighbors(u)) {
                const weight_t c = hc::relax_host(dist[nei.v], nei.w);
                if (c < dist[u]) {
                    dist[u] = c;
                    changed = true;
                }
            }
    }
    *ret_dist = dist;
}

// Dijkstra with a binary heap of (distance, vertex), stale entries skipped.
inline void sssp_dijkstras_cpu_serial(const CSRWGraph &g, nid_t source_id, weight_t **ret_dist) {
    weight_t *dist = new weight_t[g.num_nodes];
    std::fill(dist, dist + g.num_nodes, INF_WEIGHT);
    dist[source_id] = 0;
    using item = std::pair<weight_t, nid_t>;
    std::priority_queue<item, std::vector<item>, std::greater<item>> heap;
    heap.push({0, source_id});
    while (!heap.empty()) {
        const item top = heap.top();
        heap.pop();
        if (top.first != dist[top.second]) continue;
        for (wnode_t nei : g.get_neighbors(top.second)) {
            const weight_t c = hc::relax_host(top.first, nei.w);
            if (c < dist[nei.v]) {
                dist[nei.v] = c;
                heap.push({c, nei.v});
            }
        }
    }
    *ret_dist = dist;
}

// ---- PageRank on the host cores (src/kernels/cpu/pr.cuh:33-65, 157-187) ----

// epoch_pr_pull_cpu_one_to_one: vertices tid, tid+num_threads, ... of
// [start_id, end_id) take base + 0.85 * sum score[v] / deg(v) in place
// (other threads' updates of this epoch may be seen), but only when that
// moves them by more than 5e-8; *updated counts the vertices that moved.
inline void epoch_pr_pull_cpu_one_to_one(const CSRWGraph &g, float *score, const nid_t start_id,
                                         const nid_t end_id, const int tid,
                                         const int num_threads, nid_t &updated) {
    const float damp = 0.85f;
    const float base = (1.0f - damp) / (float)g.num_nodes;
    nid_t moved = 0;
    for (nid_t u = start_id + tid; u < end_id; u += num_threads) {
        float in = 0.0f;
        for (wnode_t nei : g.get_neighbors(u)) {
            float x;
            __atomic_load(&score[nei.v], &x, __ATOMIC_RELAXED);
            in += x / (float)g.get_degree(nei.v);
        }
        float s = base + damp * in;
        if (std::fabs(s - score[u]) > 0.00000005f) {
            __atomic_store(&score[u], &s, __ATOMIC_RELAXED);
            moved++;
        }
    }
    if (moved) __atomic_fetch_add(&updated, moved, __ATOMIC_RELAXED);
}

inline pr_cpu_epoch_func get_kernel(PRCPU k) {
    return k == PRCPU::one_to_one ? epoch_pr_pull_cpu_one_to_one : nullptr;
}

// pr_pull_cpu: scores start at 1/N (the reference ignores init_scores too,
// :38-42) and the epoch runs on every thread until one moves no vertex.
inline double pr_pull_cpu(const CSRWGraph &g, pr_cpu_epoch_func epoch_kernel,
                          const float *init_scores, float **const ret_score) {
    (void)init_scores;
    float *score = new float[g.num_nodes];
    const float init = 1.0f / (float)g.num_nodes;
#pragma omp parallel for
    for (nid_t i = 0; i < g.num_nodes; i++) score[i] = init;
    nid_t updated = 1;
    const double t0 = hc::now_ms();
    while (updated != 0) {
        updated = 0;
#pragma omp parallel
        epoch_kernel(g, score, 0, g.num_nodes, omp_get_thread_num(), omp_get_num_threads(),
                     updated);
    }
    const double ms = hc::now_ms() - t0;
    *ret_score = score;
    return ms;
}
inline double pr_pull_cpu(const CSRWGraph &g, PRCPU epoch_kernel, const float *init_scores,
                          float **const ret_score) {
    return pr_pull_cpu(g, get_kernel(epoch_kernel), init_scores, ret_score);
}

// PageRank on the host cores with the library's Jacobi arithmetic (FP32
// Kahan sums, FP64 1/deg, FP64 L1 residual < 1e-4, <= 20 iterations): the
// host counterpart of pr_gpu.
inline double pr_jacobi_cpu(const CSRWGraph &g, float **const ret_score) {
    float *score = new float[g.num_nodes];
    hc_run_stats_t st{};
    hc::run_cpu_schedule(g, HC_K_PR_CPU_ONE_TO_ONE, &st,
                         [](hc_hetero_t h, hc_run_stats_t *s, void *p) {
                             hc_pr_opts_t o{0.85f, 1e-4, 20, 0};
                             hc::check_or_exit(hc_pr_heterogeneous(h, &o, static_cast<float *>(p), s),
                                               "pr_jacobi_cpu");
                         },
                         score);
    *ret_score = score;
    return st.ms;
}

// ---- BFS on the host cores (src/kernels/cpu/bfs.cuh:125-397) ----

// reset_parents: every parent INVALID_NODE, the source its own parent.
inline void reset_parents(nid_t *const parents, const nid_t num_nodes, const nid_t source_id) {
#pragma omp parallel for
    for (nid_t i = 0; i < num_nodes; i++) parents[i] = INVALID_NODE;
    parents[source_id] = source_id;
}

namespace hc {
// Claims v for parent u: true for exactly one caller while v is unvisited.
inline bool claim_parent(nid_t *parents, nid_t v, nid_t u) {
    nid_t expect = INVALID_NODE;
    return __atomic_load_n(&parents[v], __ATOMIC_RELAXED) == INVALID_NODE &&
           __atomic_compare_exchange_n(&parents[v], &expect, u, false, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED);
}
}  // namespace hc

// epoch_bfs_push_cpu_by_node: frontier vertices split between threads; each
// claimed neighbour joins the next window and adds its degree to num_edges.
inline void epoch_bfs_push_cpu_by_node(const CSRUWGraph &g, nid_t *const parents,
                                       SlidingWindow<nid_t> &frontier, nid_t &num_edges) {
    const nid_t *f = frontier.begin();
    const int64_t fn = (int64_t)frontier.size();
    int64_t edges = 0;
#pragma omp parallel reduction(+ : edges)
    {
        LocalWindow<nid_t> next(frontier);
#pragma omp for schedule(dynamic, 64) nowait
        for (int64_t i = 0; i < fn; i++) {
            const nid_t u = f[i];
            for (nid_t v : g.get_neighbors(u))
                if (hc::claim_parent(parents, v, u)) {
                    next.push_back(v);
                    edges += g.get_degree(v);
                }
        }
        next.flush();
    }
    num_edges += (nid_t)edges;
}

// epoch_bfs_push_cpu_by_edge: each frontier vertex's edges split between
// threads.
inline void epoch_bfs_push_cpu_by_edge(const CSRUWGraph &g, nid_t *const parents,
                                       SlidingWindow<nid_t> &frontier, nid_t &num_edges) {
    const nid_t *f = frontier.begin();
    const int64_t fn = (int64_t)frontier.size();
    int64_t edges = 0;
#pragma omp parallel reduction(+ : edges)
    {
        LocalWindow<nid_t> next(frontier);
        for (int64_t i = 0; i < fn; i++) {
            const nid_t u = f[i];
            const offset_t b = g.index[u], e = g.index[u + 1];
#pragma omp for schedule(static) nowait
            for (offset_t k = b; k < e; k++) {
                const nid_t v = g.neighbors[k];
                if (hc::claim_parent(parents, v, u)) {
                    next.push_back(v);
                    edges += g.get_degree(v);
                }
            }
        }
        next.flush();
    }
    num_edges += (nid_t)edges;
}

// epoch_bfs_pull_cpu: every unvisited vertex of [start_id, end_id) takes
// its first neighbour (CSR order) in `frontier` as parent and joins
// next_frontier; num_nodes counts them.
inline void epoch_bfs_pull_cpu(const CSRUWGraph &g, nid_t *const parents, const nid_t start_id,
                               const nid_t end_id, const Bitmap::Bitmap *const frontier,
                               Bitmap::Bitmap *const next_frontier, nid_t &num_nodes) {
    nid_t found = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : found)
    for (nid_t u = start_id; u < end_id; u++) {
        if (parents[u] != INVALID_NODE) continue;
        for (nid_t v : g.get_neighbors(u))
            if (Bitmap::get_bit(frontier, v)) {
                parents[u] = v;
                Bitmap::set_bit_atomic(next_frontier, u);
                found++;
                break;
            }
    }
    num_nodes += found;
}

inline bfs_cpu_push_epoch_func get_kernel(BFSCPUPush k) {
    switch (k) {
        case BFSCPUPush::by_node: return epoch_bfs_push_cpu_by_node;
        case BFSCPUPush::by_edge: return epoch_bfs_push_cpu_by_edge;
        default: return nullptr;
    }
}
inline bfs_cpu_pull_epoch_func get_kernel(BFSCPUPull k) {
    return k == BFSCPUPull::pull ? epoch_bfs_pull_cpu : nullptr;
}

// conv_window_to_bitmap: the current window's vertices set in `bitmap`.
inline void conv_window_to_bitmap(const SlidingWindow<nid_t> &window,
                                  Bitmap::Bitmap *const bitmap) {
    const nid_t *w = window.begin();
    const int64_t wn = (int64_t)window.size();
#pragma omp parallel for
    for (int64_t i = 0; i < wn; i++) Bitmap::set_bit_atomic(bitmap, w[i]);
}

// conv_bitmap_to_window: the bitmap's vertices become the current window.
inline void conv_bitmap_to_window(const Bitmap::Bitmap *const bitmap,
                                  SlidingWindow<nid_t> &window, const nid_t num_nodes) {
#pragma omp parallel
    {
        LocalWindow<nid_t> local(window);
#pragma omp for nowait
        for (nid_t u = 0; u < num_nodes; u++)
            if (Bitmap::get_bit(bitmap, u)) local.push_back(u);
        local.flush();
    }
    window.slide_window();
}

// bfs_push_cpu<epoch>: top-down epochs until the frontier is empty.
template <bfs_cpu_push_epoch_func push_epoch_func>
inline double bfs_push_cpu(const CSRUWGraph &g, const nid_t source_id,
                           nid_t **const ret_parents) {
    nid_t *parents = new nid_t[g.num_nodes];
    reset_parents(parents, g.num_nodes, source_id);
    SlidingWindow<nid_t> frontier(g.num_nodes);
    frontier.push_back(source_id);
    frontier.slide_window();
    const double t0 = hc::now_ms();
    while (!frontier.empty()) {
        nid_t num_edges = 0;
        push_epoch_func(g, parents, frontier, num_edges);
        frontier.slide_window();
    }
    const double ms = hc::now_ms() - t0;
    *ret_parents = parents;
    return ms;
}

// bfs_pull_cpu: bottom-up epochs until one adds no vertex.
inline double bfs_pull_cpu(const CSRUWGraph &g, const nid_t source_id,
                           nid_t **const ret_parents) {
    nid_t *parents = new nid_t[g.num_nodes];
    reset_parents(parents, g.num_nodes, source_id);
    Bitmap::Bitmap *frontier = Bitmap::constructor(g.num_nodes);
    Bitmap::Bitmap *next = Bitmap::constructor(g.num_nodes);
    Bitmap::set_bit(frontier, source_id);
    nid_t num_nodes = 0;
    const double t0 = hc::now_ms();
    do {
        num_nodes = 0;
        epoch_bfs_pull_cpu(g, parents, 0, g.num_nodes, frontier, next, num_nodes);
        std::swap(frontier, next);
        Bitmap::reset(next);
    } while (num_nodes != 0);
    const double ms = hc::now_ms() - t0;
    Bitmap::destructor(&frontier);
    Bitmap::destructor(&next);
    *ret_parents = parents;
    return ms;
}

// ---------------------------------------------------------------------------
// Heterogeneous drivers: a segment schedule (.skd text, hc_sched.h) instead
// of a generated driver.  schedule_path NULL: the HC_SCHEDULE environment
// variable, else every segment on the GPU.
// ---------------------------------------------------------------------------

namespace hc {

inline hc_schedule_t load_schedule_or_default(const char *path, int32_t gpu_kernel) {
    if (!path) path = std::getenv("HC_SCHEDULE");
    hc_schedule_t s = nullptr;
    if (path) {
        check_or_exit(hc_schedule_load(path, &s), "hc_schedule_load");
    } else {
        hc_schedule_block_t b{HC_DEVICE_AMD_MI355X, 0, 0, 0, gpu_kernel, 0, 0};
        check_or_exit(hc_schedule_create(&b, 1, 1, &s), "hc_schedule_create");
    }
    return s;
}

template <typename NeighborT, typename Fn>
inline double run_schedule(const GenericCSRGraph<NeighborT> &g, const char *path,
                           int32_t gpu_kernel, Fn fn) {
    check_or_exit(hc_init(0), "hc_init");
    hc_schedule_t s = load_schedule_or_default(path, gpu_kernel);
    HostView<NeighborT> v(g);
    hc_hetero_t h = nullptr;
    check_or_exit(hc_hetero_create(&v.c, s, &h), "hc_hetero_create");
    hc_run_stats_t st{};
    fn(h, &st);
    hc_hetero_free(h);
    hc_schedule_free(s);
    return st.ms;
}

}  // namespace hc

// enable_all_peer_access / enable_all_peer_access_pr (generated beside the
// heterogeneous drivers, src/kernels/heterogeneous/{sssp,pr}.cuh): peer
// access between the GPUs of a schedule.  Every GPU block of a schedule runs
// on the one GPU of this process, so there is nothing to enable.
inline void enable_all_peer_access() {}
inline void enable_all_peer_access_pr() {}

// sssp_pull_heterogeneous (src/kernels/heterogeneous/sssp.cuh:39-387).
inline double sssp_pull_heterogeneous(const CSRWGraph &g, const weight_t *init_dist,
                                      weight_t **const ret_dist,
                                      const char *schedule_path = nullptr) {
    weight_t *dist = new weight_t[g.num_nodes];
    const double ms = hc::run_schedule(g, schedule_path, HC_K_SSSP_GPU_BUCKETED,
                                       [&](hc_hetero_t h, hc_run_stats_t *st) {
                                           hc::check_or_exit(hc_sssp_heterogeneous(h, init_dist, dist, st),
                                                             "sssp_pull_heterogeneous");
                                       });
    *ret_dist = dist;
    return ms;
}

// pr_pull_heterogeneous (src/kernels/heterogeneous/pr.cuh:39-346).
inline double pr_pull_heterogeneous(const CSRWGraph &g, const float *init_score,
                                    float **const ret_score,
                                    const char *schedule_path = nullptr) {
    (void)init_score;
    float *score = new float[g.num_nodes];
    const double ms = hc::run_schedule(g, schedule_path, HC_K_PR_GPU_BUCKETED,
                                       [&](hc_hetero_t h, hc_run_stats_t *st) {
                                           hc_pr_opts_t o{0.85f, 1e-4, 20};
                                           hc::check_or_exit(hc_pr_heterogeneous(h, &o, score, st),
                                                             "pr_pull_heterogeneous");
                                       });
    *ret_score = score;
    return ms;
}

// bfs_do_heterogeneous (src/kernels/heterogeneous/bfs.cuh:19-223).
inline double bfs_do_heterogeneous(const CSRUWGraph &g, const nid_t source_id,
                                   nid_t **const ret_parents,
                                   const char *schedule_path = nullptr) {
    nid_t *parents = new nid_t[g.num_nodes];
    const double ms = hc::run_schedule(g, schedule_path, HC_K_BFS_GPU_BUCKETED,
                                       [&](hc_hetero_t h, hc_run_stats_t *st) {
                                           hc::check_or_exit(hc_bfs_heterogeneous(h, source_id, 15, 18,
                                                                                  parents, nullptr, st),
                                                             "bfs_do_heterogeneous");
                                       });
    *ret_parents = parents;
    return ms;
}

#endif  // HC_CPP_KERNELS_H
