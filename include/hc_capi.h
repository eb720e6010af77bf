/*
 * hc_capi.h -- C ABI of the MI355X (gfx950) graph epoch library.
 *
 * Every entry point returns an int status (HC_OK == 0, negative on error) and
 * never calls exit(); the message of the last error on the calling thread is
 * returned by hc_last_error().  No C++ or HIP types appear in the signatures:
 * device memory is passed as plain pointers, streams as opaque hc_stream_t
 * (a hipStream_t; NULL selects the library's own stream).
 *
 * Reference boundary this replaces (chrisliu/hetero-compute):
 *   - epoch-kernel signatures            src/kernels/kernel_types.cuh:32-63
 *   - single-device drivers              src/kernels/gpu/sssp.cuh:42-100,
 *                                        src/kernels/gpu/bfs.cuh:21-95,
 *                                        src/kernels/gpu/pr.cuh:42-118,
 *                                        src/kernels/cpu/bfs.cuh:48-114
 *   - CSR container / loader / H2D copy  src/graph.cuh:26-100, 207-286
 *   - segmentation and source picking    src/graph.cuh:106-201
 *   - error convention CUDA_ERRCHK/CONDCHK (print + exit) src/cuda.cuh:14-21,
 *     src/util.h:16-22 -> status codes here.
 */
#ifndef HC_CAPI_H
#define HC_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Types and constants                                                      */
/* ------------------------------------------------------------------------ */

/* Integer distance sentinel.  The reference uses float infinity
 * (src/graph.cuh:32); with int32 weights an explicit sentinel is needed
 * (numeric_limits<int>::infinity() is 0).  Candidate distances saturate at
 * HC_INF_DIST, so INF + w stays INF. */
#define HC_INF_DIST 0x3FFFFFFF
/* src/graph.cuh:35 */
#define HC_INVALID_NODE (-1)

enum hc_status {
    HC_OK = 0,
    HC_ERR_INVALID = -1,     /* bad argument / malformed input          */
    HC_ERR_HIP = -2,         /* HIP runtime error                       */
    HC_ERR_OOM = -3,         /* device or host allocation failed        */
    HC_ERR_IO = -4,          /* file open/read/write failure            */
    HC_ERR_UNSUPPORTED = -5, /* valid request this build cannot serve   */
    HC_ERR_BUSY = -6,        /* admission refused (serving)             */
    HC_ERR_NOT_FOUND = -7,   /* unknown name / id                       */
    HC_ERR_NO_DEVICE = -8    /* no usable GPU                           */
};

typedef void *hc_stream_t;              /* hipStream_t                      */
typedef struct hc_graph_s *hc_graph_t;  /* device-resident CSR (HBM)        */

/* Vertex-to-lane mappings of an epoch kernel.  THREAD/WAVE/BLOCK are the
 * reference's one_to_one / warp_red / block_red (src/kernels/gpu/sssp.cuh:
 * 120-281) re-thought for 64-lane wavefronts; BUCKETED is the degree-bucketed
 * plan (hub chunks / workgroup / wavefront / lane-group / thread per vertex)
 * computed from the degree-sorted vertex order. */
enum hc_mapping {
    HC_MAP_THREAD = 0,
    HC_MAP_WAVE = 1,
    HC_MAP_BLOCK = 2,
    HC_MAP_BUCKETED = 3
};

/* Host-resident CSR.  index has num_nodes+1 int64 entries (reference
 * offset_t, src/graph.cuh:28); neighbors/weights have num_edges int32 entries
 * (structure-of-arrays instead of the reference's 8-byte wnode_t AoS,
 * src/graph.cuh:38-41).  weights may be NULL for an unweighted graph. */
typedef struct hc_host_graph {
    int32_t num_nodes;
    int64_t num_edges;
    int64_t *index;
    int32_t *neighbors;
    int32_t *weights;
} hc_host_graph_t;

typedef struct hc_graph_info {
    int32_t num_nodes;     /* vertices of the whole graph                    */
    int64_t num_edges;     /* directed edges held by this handle             */
    int32_t start_id;      /* first row held (range handles)                 */
    int32_t end_id;        /* one past the last row held                     */
    int32_t degree_sorted; /* 1 if degrees are non-increasing over rows      */
    int32_t weighted;
    int32_t offset_bytes;  /* 4 or 8: width of the device row offsets        */
    int32_t num_nonzero;   /* rows [start_id, start_id+num_nonzero) have deg>0
                              (sorted graphs); == rows otherwise             */
    int64_t max_degree;
    size_t device_bytes;   /* HBM held by the CSR arrays                     */
} hc_graph_info_t;

typedef struct hc_device_info {
    char name[128];
    char arch[64];
    int32_t compute_units;
    int32_t wavefront_size;
    size_t hbm_bytes;
    int32_t lds_bytes_per_block;
} hc_device_info_t;

/* Statistics of a full traversal / iteration run. */
typedef struct hc_run_stats {
    double ms;             /* time of the epoch loop.  Device-driven loops
                              (hc_bfs, hc_sssp, hc_pagerank, hc_sssp_multi):
                              the device's constant-rate wall clock, read by
                              the init kernel and by the controller of the
                              last epoch.  Host-driven loops (hc_sssp_pull,
                              hc_pr_pull, hc_bfs_many, hc_bfs_multi): HIP
                              events around the launches
                              (hc_bfs_many_noalloc: host clock).
                              Heterogeneous runs: host clock, copies
                              included.                                    */
    int32_t epochs;        /* epochs executed (with buckets, hc_sssp and
                              hc_sssp_multi count their promotion epochs
                              here too: epochs - push - pull of them)        */
    int32_t push_epochs;   /* top-down / push epochs                          */
    int32_t pull_epochs;   /* bottom-up / pull epochs                         */
    int32_t launches;      /* kernel launches issued for the loop             */
    int64_t reached;       /* vertices reached (BFS), -1 if not tracked       */
    double residual;       /* PageRank: final L1 residual (FP64)              */
    double alg_bytes;      /* algorithmic HBM bytes of the epochs run          */
    double kernel_ms;      /* with profile.kernel_events: summed HIP-event
                              durations of the main kernel's launches (the
                              step kernels of the epochs run), else 0          */
    int32_t kernel_launches; /* launches summed into kernel_ms                */
    int32_t reserved;
} hc_run_stats_t;

/* ------------------------------------------------------------------------ */
/* Runtime                                                                  */
/* ------------------------------------------------------------------------ */

const char *hc_last_error(void);
int hc_version(void);
/* Select the device and cache its properties (CU count for grid sizing). */
int hc_init(int device);
int hc_get_device_info(hc_device_info_t *out);
int hc_device_synchronize(void);
/* Free and total device memory of the current device (hipMemGetInfo). */
int hc_mem_get_info(size_t *free_bytes, size_t *total_bytes);
int hc_dmalloc(void **ptr, size_t bytes);
int hc_dfree(void *ptr);
int hc_hmalloc_pinned(void **ptr, size_t bytes);
int hc_hfree_pinned(void *ptr);
/* kind: 1 host->device, 2 device->host, 3 device->device */
int hc_memcpy(void *dst, const void *src, size_t bytes, int kind,
              hc_stream_t stream);
int hc_memset_d32(void *dst, int32_t value, size_t count, hc_stream_t stream);
int hc_stream_create(hc_stream_t *stream);
int hc_stream_destroy(hc_stream_t stream);
int hc_stream_synchronize(hc_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Host graphs: generator/builder, file I/O, sources, segments              */
/* ------------------------------------------------------------------------ */

/* gapbs-style synthetic graph (src/gapbs/generator.h:34-172,
 * src/gapbs/builder.h:39-407, src/gapbs/gapbs.cuh:121-125): R-MAT
 * (A=.57,B=.19,C=.19) with an ID permutation, or uniform; 2^18-edge blocks
 * reseeded with seed+block; symmetrized; self-loops and duplicate edges
 * removed (the lightest duplicate is kept); relabelled by descending degree
 * (ties: larger old id first); neighbour lists sorted.  Weighted graphs get
 * int32 weights uniform in [1,255], equal in both directions. */
int hc_host_graph_generate(int uniform, int scale, int degree, int weighted,
                           int64_t seed, hc_host_graph_t *out);
/* Build a CSR from an edge list with the same builder steps. */
int hc_host_graph_from_edges(const int32_t *src, const int32_t *dst,
                             const int32_t *weights, int64_t num_edges,
                             int32_t num_nodes, int symmetrize, int relabel,
                             hc_host_graph_t *out);
/* Reference binary format (src/graph.cuh:207-239, src/gapbs/gapbs.cuh:
 * 58-112): int32 N, int64 m, int64 index[N+1], then m 8-byte {int32 v;
 * float w} (.wsg, weighted=1) or m int32 ids (.sg, weighted=0).  Weights read
 * from a .wsg must be integral and in [0, 2^30). */
int hc_host_graph_read(const char *path, int weighted, hc_host_graph_t *out);
int hc_host_graph_write(const char *path, const hc_host_graph_t *g,
                        int weighted);
void hc_host_graph_free(hc_host_graph_t *g);
/* SourcePicker (src/graph.cuh:106-141): mt19937(seed), uniform ids, rejects
 * degree-0 vertices and lone self-loops.  seed 37525 is the reference's. */
int hc_pick_sources(const hc_host_graph_t *g, int64_t seed, int32_t count,
                    int32_t *out);
/* compute_equal_edge_ranges (src/graph.cuh:156-201): out has
 * num_segments+1 entries. */
int hc_equal_edge_ranges(const hc_host_graph_t *g, int32_t num_segments,
                         int32_t chunk_size, int32_t *out);

/* ------------------------------------------------------------------------ */
/* Device graphs (CSR in HBM)                                               */
/* ------------------------------------------------------------------------ */

/* Whole graph.  Offsets are stored as uint32 when the edge count allows,
 * otherwise uint64; ids and weights as int32 structure-of-arrays. */
int hc_graph_upload(const hc_host_graph_t *g, hc_graph_t *out);
/* Rows [start_id, end_id) only, with down-shifted offsets
 * (copy_subgraph_to_device, src/graph.cuh:250-286).  Epoch calls on a range
 * handle take global vertex ids; per-vertex state stays global-length. */
int hc_graph_upload_range(const hc_host_graph_t *g, int32_t start_id,
                          int32_t end_id, hc_graph_t *out);
int hc_graph_free(hc_graph_t g);
/* A second handle over the same device CSR with its own plans and
 * workspaces, so that runs on different streams can proceed concurrently
 * (runs on one handle are serialised).  g must outlive the share; freeing
 * the share leaves g's arrays in place.  No reference counterpart. */
int hc_graph_share(hc_graph_t g, hc_graph_t *out);
int hc_graph_get_info(hc_graph_t g, hc_graph_info_t *out);

/* Tuning knobs, process-wide (the autotuner's interface; no reference
 * counterpart -- the reference fixed these per generated driver,
 * src/kernels/heterogeneous/sssp.cuh:180-278).  Keys:
 *   grid_per_cu            workgroups per CU of full-device launches (8); the
 *                          loops launch at most as many as fit at once
 *   sssp.alpha             SSSP pull when frontier edges > m/alpha; 0: by graph
 *                          shape, 4 when the top 1% rows hold >= 5% of the
 *                          edges (skewed), else 2 (0)
 *   sssp.grid_per_cu       cap on workgroups per CU of the SSSP loop (4)
 *   sssp.filter            1: SSSP pull epochs gather only neighbours whose
 *                          distance changed in the previous epoch; k >= 2:
 *                          only when fewer than n / k vertices changed
 *                          (column-blocked pulls never filter); 0: never (0)
 *   sssp.dist16            1: SSSP pull epochs gather a 16-bit mirror of the
 *                          distances (rebuilt before each pull; exact); 2: on
 *                          for flat graphs only (2)
 *   sssp.stage             distances staged in LDS by SSSP pull sweeps (0)
 *   sssp.block_threshold   SSSP: degree above which a workgroup takes a row
 *                          (4096; the hub class is pull.*'s)
 *   sssp.wave_threshold    SSSP: degree above which a wavefront takes a
 *                          row (96)
 *   sssp.g16_threshold / sssp.g4_threshold   SSSP: degree above which 16 / 4
 *                          lanes take a row (0 / 8; 0: 24)
 *   pull.hub_threshold     degree above which a row is split into chunks
 *   pull.hub_chunk         edges per chunk (>= 256)
 *   pull.block_threshold   degree above which a workgroup takes a row
 *   pull.wave_threshold    ... a wavefront
 *   pull.g16_threshold     ... 16 lanes
 *   pull.g4_threshold      ... 4 lanes (else one lane)
 *   bu.wave_threshold / bu.g16_threshold / bu.g4_threshold   BFS bottom-up
 *   pull.stage / bu.stage  per-vertex entries / bitmap words staged in LDS
 *   push.stage             output entries a push-epoch workgroup stages in LDS
 *                          before reserving queue space (<= 15360; 8192);
 *                          past it, entries go to the queues one atomic each
 *   push.chunk             edges per workgroup chunk of a hub vertex in push
 *                          epochs (>= 128; 256)
 *   bu.stage_min           a workgroup whose unvisited list is shorter than
 *                          this reads the frontier bitmap without LDS
 *                          staging (4096)
 *   compact.grid           workgroups of the frontier compaction launch, at
 *                          most the step grid (4096)
 *   bu.first               bottom-up rows test their first neighbour from a
 *                          per-vertex array (n int32 per graph) before the
 *                          edge array: 1 one-lane rows, 2 all rows, 0 off (2)
 *   sssp.push_dedup        SSSP push epochs deduplicate the vertices they lower
 *                          while staging them in LDS (1: one returning atomic
 *                          per lowering edge) or on the spot (0: two); 2
 *                          (default): 1 on flat graphs, 0 on skewed ones,
 *                          where hubs' neighbours repeat and fill the staging
 *   bu.remaining           a bottom-up epoch after another sweeps only the
 *                          vertices the previous one left unvisited, with
 *                          this many lanes per vertex (1, 4 or 16; 0: off)
 *   loop.batch             epochs the host enqueues between two reads of the
 *                          loop's done flag (1..16; 2): more hides more host
 *                          latency, and up to twice as many no-op epochs
 *                          are queued past the end of a loop
 *   profile.kernel_events  1: time each launch of the loop's main kernel
 *                          (hc_run_stats_t.kernel_ms); 0 by default
 *   multi.graph            1: batched loops relaunch an epoch hipGraph
 *   multi.grid_per_cu      workgroups per CU of the batched sweep (16)
 *   multi.block            threads per workgroup of the batched sweep (128,
 *                          256, 512 or 1024; 256)
 *   multi.delta            batched SSSP: a distance improvement propagates at
 *                          once only while below a threshold that rises by
 *                          this much whenever nothing below it changes
 *                          (delta-stepping buckets over the 64 sources);
 *                          0 off; -1 (default) 12 x the graph's sampled
 *                          largest weight / mean degree on flat graphs, off
 *                          on skewed ones.  Results do not depend on it
 *   sssp.delta             hc_sssp: a vertex lowered to a value at or above
 *                          a threshold waits in a far set instead of the next
 *                          frontier; when a frontier comes out empty the
 *                          threshold rises by this much (doubling after a
 *                          promotion that found nothing) and the far
 *                          vertices below it become the frontier (near-far
 *                          buckets; push epochs only, no pulls); 0 off; -1
 *                          (default) 2 x the graph's sampled largest weight
 *                          / mean degree on flat graphs, off on skewed ones
 *                          and with direction 1.  Results do not depend on it
 *   multi.push_div         batched sweep: an epoch whose flagged vertices
 *                          hold at most m / this edges (none more than
 *                          1024) pushes from them with atomicMin instead of
 *                          sweeping every row (16; 0 never)
 *   msbfs.on               hc_bfs_multi runs bit-parallel levels (1, default)
 *                          or the batched sweep with unit weights (0)
 *   msbfs.alpha            ... a bit-parallel level pushes while its frontier
 *                          holds fewer than m / this edges, else pulls (32)
 *   msbfs.col_shift        ... its pulls gather column blocks of 2^this
 *                          vertices (0 off; -1 default: four blocks on flat
 *                          graphs, sixteen on skewed ones, >= 2^16 each)
 *   multi.hub_threshold / multi.hub_chunk / multi.wave_threshold /
 *   multi.g16_threshold    degree classes of the batched sweep
 *   pull.col_shift         pull sweeps of whole graphs with at least two
 *                          blocks of 2^this vertices gather block by block
 *                          from a column-blocked copy of the CSR (built on
 *                          first use: as many bytes again as ids + weights,
 *                          plus K (n+1) offsets; at most 32 blocks); 0 off,
 *                          6..30 on every graph, -1 (default) automatic:
 *                          flat graphs only, four blocks of >= 2^20 vertices
 *   cb.min_slice           rows of degree above this x blocks are swept block
 *                          by block, lighter rows once over the plain CSR (4)
 *   cb.g4_slice / cb.g16_slice / cb.wave_slice / cb.block_slice   mean
 *                          slice degree (degree / blocks) above which 4 / 16
 *                          lanes, a wavefront, a workgroup take a row's slice
 *                          in column-blocked sweeps (2 / 12 / 48 / 512)
 *   many.block             threads per workgroup of hc_bfs_many (256, 512
 *                          or 1024; 512)
 *   many.per_cu            its workgroups per CU (at most one per graph; 0:
 *                          as many as the device holds with the launch's
 *                          LDS; 2)
 *   many.light_deg         its top-down levels: a thread expands frontier
 *                          vertices up to this degree (32) ...
 *   many.heavy_cap         ... heavier ones, up to this many per level (at
 *                          most 1024; 512), by whole wavefronts
 *   many.lds_nodes         graphs of up to this many vertices (at most 2^17)
 *                          keep their bitmaps in LDS, sized to the largest
 *                          of them in the batch; larger ones use global
 *                          bitmaps (2^17)
 * Plans are rebuilt for new values; changing knobs while another thread
 * runs on the same graph is not supported. */
int hc_set_tuning(const char *key, int64_t value);
int hc_get_tuning(const char *key, int64_t *value);
int hc_reset_tuning(void);

/* Knobs of one graph (no counterpart in the reference, whose tuned choices
 * are baked into generated drivers per graph): hc_graph_set_tuning records
 * key = value on g and every share of it (made before or after); each call
 * on those handles then runs with the process-wide knobs and the graph's
 * recorded ones applied over them, so calls on different graphs in other
 * threads (a server's slots) use their own.  Values are checked as by
 * hc_set_tuning.  hc_graph_get_tuning reads the graph's value, else the
 * process-wide one; hc_graph_reset_tuning drops the graph's. */
int hc_graph_set_tuning(hc_graph_t g, const char *key, int64_t value);
int hc_graph_get_tuning(hc_graph_t g, const char *key, int64_t *value);
int hc_graph_reset_tuning(hc_graph_t g);

/* ------------------------------------------------------------------------ */
/* Epoch kernels (one bulk-synchronous sweep of [start_id, end_id))         */
/* ------------------------------------------------------------------------ */
/* All state pointers are device pointers of global length.  Counters are
 * accumulated (add), so the caller zeroes them, as the reference drivers do
 * (src/kernels/gpu/sssp.cuh:78).  blocks/threads apply to THREAD/WAVE/BLOCK
 * (threads % 64 == 0 for WAVE and BLOCK; 0 selects a default); BUCKETED
 * ignores them. */

/* epoch_sssp_pull_gpu_{one_to_one,warp_red,block_red}
 * (src/kernels/gpu/sssp.cuh:120-281; signature kernel_types.cuh:36-37):
 * dist[u] = min(dist[u], min_{(v,w) in N(u)} dist[v] + w) in place;
 * *updated += number of vertices whose distance changed. */
int hc_sssp_pull_epoch(hc_graph_t g, int mapping, int32_t start_id,
                       int32_t end_id, int32_t *dist, int32_t *updated,
                       int blocks, int threads, hc_stream_t stream);

/* epoch_bfs_pull_gpu_* (src/kernels/gpu/bfs.cuh:116-288, signature
 * kernel_types.cuh:60-63) / epoch_bfs_pull_cpu (src/kernels/cpu/bfs.cuh:
 * 305-335): every unvisited u (parents[u] == -1) whose neighbour list holds a
 * vertex of the frontier bitmap takes the first such neighbour (CSR order) as
 * parent, is set in next_frontier, and counts into *num_nodes.  Bitmaps are
 * 32-bit words, bit (v & 31) of word v >> 5 (src/bitmap.cuh:19-34). */
int hc_bfs_pull_epoch(hc_graph_t g, int mapping, int32_t start_id,
                      int32_t end_id, int32_t *parents,
                      const uint32_t *frontier, uint32_t *next_frontier,
                      int32_t *num_nodes, int blocks, int threads,
                      hc_stream_t stream);

/* epoch_bfs_push_cpu_by_node (src/kernels/cpu/bfs.cuh:212-244), on the GPU:
 * for each u of frontier[0..frontier_size), each unvisited neighbour v gets
 * parents[v] = u (compare-and-swap from -1), is appended to next_frontier
 * (*next_size += 1) and adds deg(v) to *num_edges (the scout count). */
int hc_bfs_push_epoch(hc_graph_t g, const int32_t *frontier,
                      int32_t frontier_size, int32_t *parents,
                      int32_t *next_frontier, int32_t *next_size,
                      int64_t *num_edges, hc_stream_t stream);

/* PageRank pull epoch, Jacobi form of epoch_pr_pull_{cpu,gpu}_one_to_one
 * (src/kernels/cpu/pr.cuh:157-187, src/kernels/gpu/pr.cuh:140-174):
 * score_out[u] = (1-d)/N + d * sum_{v in N(u)} contrib[v], the sum in FP32
 * with Kahan compensation, and contrib_out[u] = (float)(score_out[u] *
 * inv_degree[u]) with inv_degree in FP64.  residual_partials receives one
 * FP64 partial of sum |score_out - score_in| per workgroup (may be NULL). */
int hc_pr_pull_epoch(hc_graph_t g, int32_t start_id, int32_t end_id,
                     float damping, const float *score_in,
                     const float *contrib, const double *inv_degree,
                     float *score_out, float *contrib_out,
                     double *residual_out, hc_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Full runs with the graph resident (timer scope = epoch loop only,        */
/* src/kernels/gpu/sssp.cuh:76-86)                                          */
/* ------------------------------------------------------------------------ */

typedef struct hc_sssp_opts {
    int32_t direction;  /* 0 auto push/pull, 1 pull only, 2 push only        */
    int32_t alpha;      /* go pull when frontier edges > m/alpha (0 = default)*/
    int32_t max_epochs; /* 0 = unbounded                                     */
} hc_sssp_opts_t;

/* Single-source SSSP, int32 weights >= 0: dist (device, num_nodes) receives
 * the distances, HC_INF_DIST where unreachable.  Epoch loop of Bellman-Ford
 * pull sweeps (degree-bucketed) and push epochs over the active frontier,
 * chosen per epoch; the fixpoint equals sssp_pull_cpu's
 * (src/kernels/cpu/sssp.cuh:33-65). */
int hc_sssp(hc_graph_t g, int32_t source, int32_t *dist,
            const hc_sssp_opts_t *opts, hc_run_stats_t *stats,
            hc_stream_t stream);

/* sssp_pull_gpu (src/kernels/gpu/sssp.cuh:42-100): pull epochs with the given
 * mapping from init_dist (device) until no vertex changes. */
int hc_sssp_pull(hc_graph_t g, int mapping, const int32_t *init_dist,
                 int32_t *dist, int blocks, int threads,
                 hc_run_stats_t *stats, hc_stream_t stream);

typedef struct hc_bfs_opts {
    int32_t alpha;     /* push -> pull when scout > edges_to_check/alpha (15) */
    int32_t beta;      /* pull -> push when n <= prev and n <= N/beta (18)    */
    int32_t direction; /* 0 direction-optimizing, 1 pull only, 2 push only   */
} hc_bfs_opts_t;

/* bfs_do_cpu<alpha,beta> (src/kernels/cpu/bfs.cuh:48-114) on the GPU with the
 * same switching rule.  depth (device, may be NULL) gets hop counts (-1 if
 * unreached); parents (device, may be NULL) a valid BFS tree with
 * parents[source] == source. */
int hc_bfs(hc_graph_t g, int32_t source, int32_t *depth, int32_t *parents,
           const hc_bfs_opts_t *opts, hc_run_stats_t *stats,
           hc_stream_t stream);

typedef struct hc_pr_opts {
    float damping;     /* 0.85 (src/kernels/cpu/pr.cuh:139)                  */
    double tolerance;  /* stop when the FP64 L1 residual < tolerance (1e-4);
                          0 = default, < 0 = run exactly max_iters           */
    int32_t max_iters; /* 20                                                 */
    int32_t rule;      /* 0: the residual rule above.  1: pr_pull_cpu's rule
                          as written (src/kernels/cpu/pr.cuh:165-181): a
                          vertex takes its new score only when it moves by
                          more than 5e-8; stop after an iteration that moved
                          no vertex (tolerance unused).  hc_pagerank only.  */
} hc_pr_opts_t;

/* PageRank pull, Jacobi iteration from 1/N (src/kernels/cpu/pr.cuh:33-65
 * start value); scores (device, num_nodes) in FP32. */
int hc_pagerank(hc_graph_t g, float *scores, const hc_pr_opts_t *opts,
                hc_run_stats_t *stats, hc_stream_t stream);

/* pr_pull_gpu (src/kernels/gpu/pr.cuh:42-118) with its epoch kernel, start
 * scores and launch shape honoured: one epoch per iteration of the given
 * mapping (THREAD / WAVE / BLOCK: a thread, wavefront or workgroup per
 * vertex, the counterparts of epoch_pr_pull_gpu_{one_to_one,warp_red,
 * block_red}, :140-318, on `blocks` x `threads`, 0 = 64 x 1024, threads a
 * multiple of 64 for WAVE and BLOCK; BUCKETED: the degree-bucketed sweep,
 * library launch shape), the stopping measure read back after each
 * iteration as the reference's host loop does.  Same Jacobi arithmetic and
 * opts as hc_pagerank; init_score (device, may be NULL = 1/N) seeds the
 * first iteration.  stats->ms: HIP events around the loop. */
int hc_pr_pull(hc_graph_t g, int mapping, const float *init_score,
               float *scores, int blocks, int threads,
               const hc_pr_opts_t *opts, hc_run_stats_t *stats,
               hc_stream_t stream);

/* Per-epoch record of the last hc_sssp (algorithm 0) or hc_bfs (1) run on g:
 * the direction each epoch ran (0 push / top-down, 1 pull / bottom-up), the
 * size and degree sum of the frontier it produced and the edges it examined.
 * No reference counterpart; it exposes what the reference's commented-out
 * per-iteration prints showed (src/kernels/cpu/bfs.cuh:82-83, 99-100). */
typedef struct hc_epoch_trace {
    int32_t count;
    int8_t mode[128];      /* 0 push, 1 pull, 2 promotion (SSSP buckets)      */
    uint32_t frontier[128];
    uint64_t frontier_edges[128];
    uint64_t scanned[128];
    double us[128];        /* device time of the epoch (end of the previous
                              epoch's controller to end of this one's)       */
} hc_epoch_trace_t;
int hc_get_epoch_trace(hc_graph_t g, int algorithm, hc_epoch_trace_t *out);

#ifdef __cplusplus
}
#endif

#endif /* HC_CAPI_H */
