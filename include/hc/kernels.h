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
#include <ostream>
#include <string>
#include <vector>

#include "../hc_batch.h"
#include "../hc_sched.h"
#include "graph.h"

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
enum class BFSGPU { one_to_one, warp, bucketed, undefined };

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
inline int32_t kernel_id_of(BFSGPU k) {
    return k == BFSGPU::undefined ? -1 : HC_K_BFS_GPU_ONE_TO_ONE + (int32_t)k;
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

// Epoch-kernel types (src/kernels/kernel_types.cuh): CPU epochs keep the
// reference signature; GPU epochs are registry ids.
typedef void (*sssp_cpu_epoch_func)(const CSRWGraph &, weight_t *, const nid_t, const nid_t,
                                    const int, const int, nid_t &);
typedef SSSPGPU sssp_gpu_epoch_func;
typedef PRGPU pr_gpu_epoch_func;
typedef BFSGPU bfs_gpu_epoch_func;

namespace hc {

inline int mapping_of(SSSPGPU k) {
    switch (k) {
        case SSSPGPU::one_to_one: return HC_MAP_THREAD;
        case SSSPGPU::warp_red: return HC_MAP_WAVE;
        case SSSPGPU::block_red: return HC_MAP_BLOCK;
        default: return HC_MAP_BUCKETED;
    }
}
inline int mapping_of(BFSGPU k) {
    switch (k) {
        case BFSGPU::one_to_one: return HC_MAP_THREAD;
        case BFSGPU::warp: return HC_MAP_WAVE;
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

// pr_pull_gpu (src/kernels/gpu/pr.cuh:42-118), in the Jacobi form with
// FP32 Kahan sums and the FP64 residual test (L1 < 1e-4, <= 20 iterations).
// Every registry id runs the degree-bucketed sweep (the mapping changes the
// summation order only, within the stated tolerance); the launch shape is
// the library's.  Scores start at 1/N (the reference's init_score argument
// is ignored by its CPU driver and seeded with INF by its benchmarks).
template <typename GraphT>
inline double pr_pull_gpu(const GraphT &g, pr_gpu_epoch_func epoch_kernel, float *init_score,
                          float **const ret_score, int block_count = 64,
                          int thread_count = 1024) {
    (void)epoch_kernel;
    (void)init_score;
    (void)block_count;
    (void)thread_count;
    hc::DeviceGraph dg(g);
    hc::DeviceBuffer<float> score(g.num_nodes);
    hc_pr_opts_t o{0.85f, 1e-4, 20};
    hc_run_stats_t st{};
    hc::check_or_exit(hc_pagerank(dg.get(), score.p, &o, &st, nullptr), "pr_pull_gpu");
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

// pr_pull_cpu (src/kernels/cpu/pr.cuh:33-65), Jacobi form as pr_pull_gpu.
inline double pr_pull_cpu(const CSRWGraph &g, PRCPU epoch_kernel, const float *init_scores,
                          float **const ret_score) {
    (void)epoch_kernel;
    (void)init_scores;
    float *score = new float[g.num_nodes];
    hc_run_stats_t st{};
    hc::run_cpu_schedule(g, HC_K_PR_CPU_ONE_TO_ONE, &st,
                         [](hc_hetero_t h, hc_run_stats_t *s, void *p) {
                             hc_pr_opts_t o{0.85f, 1e-4, 20};
                             hc::check_or_exit(hc_pr_heterogeneous(h, &o, static_cast<float *>(p), s),
                                               "pr_pull_cpu");
                         },
                         score);
    *ret_score = score;
    return st.ms;
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
