/*
 * hc_sched.h -- segment scheduler: schedules that split the equal-edge
 * segments of a degree-sorted graph between host cores and the GPU, and the
 * runtime that executes them epoch by epoch.
 *
 * Reference boundary this replaces:
 *   - schedule file types and parser  src/schedule.h:32-226 (dead code in the
 *     reference: it includes headers that do not exist, :22-25)
 *   - device names                    src/devices.h:16-90
 *   - kernel registries / names       src/kernels/cpu/sssp.cuh:189-251,
 *     src/kernels/gpu/sssp.cuh:288-356, src/kernels/cpu/pr.cuh:193-256,
 *     src/kernels/gpu/pr.cuh:325-393, src/kernels/cpu/bfs.cuh:404-531,
 *     src/kernels/gpu/bfs.cuh:295-365
 *   - generated heterogeneous drivers src/kernels/heterogeneous/sssp.cuh:39-387,
 *     src/kernels/heterogeneous/pr.cuh:39-346,
 *     src/kernels/heterogeneous/bfs.cuh:19-223
 *   - CPU epoch signature             src/kernels/kernel_types.cuh:32-33
 */
#ifndef HC_SCHED_H
#define HC_SCHED_H

#include "hc_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

enum hc_device_type { HC_DEV_CPU = 0, HC_DEV_GPU = 1 };

/* Device names a schedule may carry: the reference's five (src/devices.h:
 * 16-21, strings :48-57) and this library's two.  All GPU devices of a
 * schedule run on the calling process's current GPU; all CPU devices on its
 * host cores. */
enum hc_device {
    HC_DEVICE_INTEL_I7_9700K = 0,       /* "Intel i7-9700K"          CPU */
    HC_DEVICE_INTEL_XEON_E5_2686 = 1,   /* "Intel Xeon E5-2686"      CPU */
    HC_DEVICE_NVIDIA_QUADRO_RTX_4000 = 2, /* "NVIDIA Quadro RTX 4000" GPU */
    HC_DEVICE_NVIDIA_TESLA_K80 = 3,     /* "NVIDIA Tesla K80"        GPU */
    HC_DEVICE_NVIDIA_TESLA_M60 = 4,     /* "NVIDIA Tesla M60"        GPU */
    HC_DEVICE_AMD_MI355X = 5,           /* "AMD Instinct MI355X"     GPU */
    HC_DEVICE_HOST_CPU = 6,             /* "Host CPU"                CPU */
    HC_DEVICE_COUNT = 7
};

/* Kernel identities.  The strings are the reference's to_string() names
 * (what .skd files carry); BUCKETED entries are this library's
 * degree-bucketed GPU sweeps. */
enum hc_kernel_id {
    HC_K_SSSP_CPU_ONE_TO_ONE = 0, /* "SSSP CPU one-to-one"  */
    HC_K_SSSP_GPU_ONE_TO_ONE = 1, /* "SSSP GPU one-to-one"  */
    HC_K_SSSP_GPU_WARP_RED = 2,   /* "SSSP GPU warp-red"    */
    HC_K_SSSP_GPU_BLOCK_RED = 3,  /* "SSSP GPU block-red"   */
    HC_K_SSSP_GPU_BUCKETED = 4,   /* "SSSP GPU bucketed"    */
    HC_K_PR_CPU_ONE_TO_ONE = 5,   /* "PR CPU one-to-one"    */
    HC_K_PR_GPU_ONE_TO_ONE = 6,   /* "PR GPU one-to-one"    */
    HC_K_PR_GPU_WARP_RED = 7,     /* "PR GPU warp-red"      */
    HC_K_PR_GPU_BLOCK_RED = 8,    /* "PR GPU block-red"     */
    HC_K_PR_GPU_BUCKETED = 9,     /* "PR GPU bucketed"      */
    HC_K_BFS_CPU_PULL = 10,       /* "BFS CPU pull"         */
    HC_K_BFS_GPU_ONE_TO_ONE = 11, /* "BFS GPU one-to-one"   */
    HC_K_BFS_GPU_WARP = 12,       /* "BFS GPU warp"         */
    HC_K_BFS_GPU_BUCKETED = 13,   /* "BFS GPU bucketed"     */
    HC_K_BFS_CPU_PUSH_BY_NODE = 14, /* "BFS CPU push by node" */
    HC_K_BFS_CPU_PUSH_BY_EDGE = 15, /* "BFS CPU push by edge" */
    HC_K_COUNT = 16
};

/* Name lookups.  hc_kernel_from_string matches by prefix, as get_kernel_id
 * does (src/schedule.h:133-142), so "SSSP GPU block-red 256 1024" names
 * HC_K_SSSP_GPU_BLOCK_RED. */
int hc_kernel_from_string(const char *name, int32_t *kernel);
const char *hc_kernel_to_string(int32_t kernel);
int hc_kernel_device_type(int32_t kernel, int32_t *device_type);
/* Exact match of a whole line against the device names (get_device,
 * src/devices.h:67-72); HC_NOT_FOUND when none matches. */
int hc_device_from_string(const char *name, int32_t *device);
const char *hc_device_to_string(int32_t device);
int hc_device_type_of(int32_t device, int32_t *device_type);

typedef struct hc_schedule_s *hc_schedule_t;

typedef struct hc_schedule_block {
    int32_t device;        /* hc_device                                    */
    int32_t device_id;     /* instance of that device name (0-based)       */
    int32_t start_segment; /* inclusive                                    */
    int32_t end_segment;   /* inclusive                                    */
    int32_t kernel;        /* hc_kernel_id                                 */
    int32_t blocks;        /* GPU launch shape (0 = library default)       */
    int32_t threads;
} hc_schedule_block_t;

/* .skd text (load_sssp_schedule, src/schedule.h:151-226; emitted by
 * write_schedule, util/scheduler/sched.py:263-286): a device-name line opens
 * a new instance of that device (ids count the appearances of each name);
 * then pairs of lines "<start_seg> <end_seg>" (inclusive) and
 * "<kernel name>[ <blocks> <threads>]".  The segment count is the largest
 * end segment + 1.  Unlike the reference parser, malformed text (a segment
 * line before any device line, a kernel of another device type, a missing
 * kernel line, start > end) is an HC_INVALID error naming the line. */
int hc_schedule_parse(const char *text, hc_schedule_t *out);
int hc_schedule_load(const char *path, hc_schedule_t *out);
int hc_schedule_create(const hc_schedule_block_t *blocks, int32_t num_blocks,
                       int32_t num_segments, hc_schedule_t *out);
int hc_schedule_info(hc_schedule_t s, int32_t *num_blocks,
                     int32_t *num_segments);
int hc_schedule_get_block(hc_schedule_t s, int32_t i, hc_schedule_block_t *out);
/* Writes .skd text into buf (cap bytes incl. NUL; buf may be NULL to size
 * it); *needed gets the size including the NUL.  One device line per device
 * instance (in order of first appearance) followed by its blocks, as
 * write_schedule does; parsing the text back gives the same instances. */
int hc_schedule_to_text(hc_schedule_t s, char *buf, size_t cap, size_t *needed);
int hc_schedule_free(hc_schedule_t s);

/* A schedule bound to a host graph: the GPU blocks' row ranges resident in
 * HBM (one range handle per block, down-shifted offsets), pinned host state
 * and one stream per GPU block.  Every segment must be covered exactly once
 * by the blocks of one algorithm. */
typedef struct hc_hetero_s *hc_hetero_t;
int hc_hetero_create(const hc_host_graph_t *g, hc_schedule_t s,
                     hc_hetero_t *out);
int hc_hetero_free(hc_hetero_t h);

/* sssp_pull_heterogeneous (src/kernels/heterogeneous/sssp.cuh:39-387): pull
 * epochs, GPU blocks on their streams while host cores run the CPU blocks
 * on pinned host state; when the schedule has CPU blocks, GPU-owned ranges
 * are copied to the host every epoch and CPU-owned ranges to the device when
 * another epoch follows (an all-GPU schedule copies the state back once, at
 * the end); stop when no vertex changed anywhere.  init_dist and dist are
 * host arrays. */
int hc_sssp_heterogeneous(hc_hetero_t h, const int32_t *init_dist,
                          int32_t *dist, hc_run_stats_t *stats);
/* pr_pull_heterogeneous (src/kernels/heterogeneous/pr.cuh:39-346) in the
 * Jacobi form of hc_pagerank: GPU and CPU blocks compute the same FP32
 * Kahan formula; the FP64 L1 residual is the sum over both. */
int hc_pr_heterogeneous(hc_hetero_t h, const hc_pr_opts_t *opts,
                        float *scores, hc_run_stats_t *stats);
/* bfs_do_heterogeneous (src/kernels/heterogeneous/bfs.cuh:19-223) with the
 * schedule's pull blocks for bottom-up epochs and host-core top-down epochs,
 * switched by the bfs_do_cpu rule (src/kernels/cpu/bfs.cuh:68-103) checked
 * after every epoch.  A "BFS CPU push by node" / "... by edge" block (what
 * the reference's push_pull_scheduler emits for push epochs) selects how the
 * top-down epochs split their work between host threads (frontier vertices
 * or each vertex's edges) and covers no segment; a schedule of push blocks
 * only pulls on the host cores.  parents and depth are host arrays (depth
 * may be NULL). */
int hc_bfs_heterogeneous(hc_hetero_t h, int32_t source, int32_t alpha,
                         int32_t beta, int32_t *parents, int32_t *depth,
                         hc_run_stats_t *stats);

/* Host-core epochs (the CPU epoch signature, src/kernels/kernel_types.cuh:
 * 32-33 and :53-57, over the C graph view): called by every thread of an
 * OpenMP team; vertex tid, tid+nthreads, ... of [start_id, end_id);
 * *updated is increased atomically. */
void hc_cpu_sssp_pull_epoch(const hc_host_graph_t *g, int32_t *dist,
                            int32_t start_id, int32_t end_id, int tid,
                            int nthreads, int32_t *updated);
/* Jacobi PageRank epoch on the host with the library's arithmetic (FP32
 * Kahan sum of contrib, FP64 inverse degree); returns this thread's FP64
 * share of the L1 residual. */
double hc_cpu_pr_pull_epoch(const hc_host_graph_t *g, float damping,
                            const float *score_in, const float *contrib,
                            const double *inv_degree, float *score_out,
                            float *contrib_out, int32_t start_id,
                            int32_t end_id, int tid, int nthreads);

/* Host-core epoch time of rows [start_id, end_id) with all threads of the
 * default OpenMP team, mean over `repeats` (algorithm 0 SSSP pull, 1 BFS
 * bottom-up, 2 PageRank), on a steady state in which every row is scanned
 * in full: the per-segment CPU measurement the schedule autotuner uses (the
 * reference measures segments with TreeBenchmark,
 * src/benchmarks/benchmark.cuh:87-309). */
int hc_cpu_epoch_ms(const hc_host_graph_t *g, int32_t algorithm, int32_t start_id,
                    int32_t end_id, int32_t repeats, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* HC_SCHED_H */
