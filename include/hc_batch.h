/*
 * hc_batch.h -- batched queries: many sources of one graph, and many small
 * graphs, each in one device-driven loop.
 *
 * No reference counterpart: the reference answers one query per driver call
 * (sssp_pull_gpu, src/kernels/gpu/sssp.cuh:42-100; bfs_gpu,
 * src/kernels/gpu/bfs.cuh:21-95), and its benchmarks loop over sources on the
 * host (src/benchmarks/gpu_benchmark.cuh:541-594).  Results are identical to
 * those single-query calls.
 */
#ifndef HC_BATCH_H
#define HC_BATCH_H

#include "hc_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HC_MULTI_WIDTH 64 /* sources per batched SSSP call */

/* SSSP from up to 64 sources at once.  sources: host array of num_sources
 * vertex ids.  dist: device array of num_nodes * 64 int32, vertex-major:
 * dist[v * 64 + q] is the distance from sources[q] to v (HC_INF_DIST when
 * unreachable; columns q >= num_sources stay HC_INF_DIST).  Every column is
 * bit-identical to hc_sssp from sources[q].  max_epochs <= 0: run to
 * convergence.  stats->ms is the device time of the whole batch. */
int hc_sssp_multi(hc_graph_t g, const int32_t *sources, int32_t num_sources,
                  int32_t *dist, int32_t max_epochs, hc_run_stats_t *stats,
                  hc_stream_t stream);

/* Breadth-first depths from up to 64 sources at once (bfs_gpu,
 * src/kernels/gpu/bfs.cuh:21-95, batched), level-synchronous over 64-bit
 * masks per vertex: the sources that reached it and, per level, those whose
 * frontier it is on; each level pushes from the frontier or pulls over all
 * rows by the frontier's edge count (knob msbfs.on = 0: the batched SSSP
 * sweep with every edge weighing 1 instead).  Weights, if any, are ignored.
 * depth[v * 64 + q] is the BFS level of v from sources[q], HC_INF_DIST when
 * unreachable (hc_bfs reports -1 there) or, with max_epochs > 0, deeper
 * than max_epochs levels; identical to hc_bfs's depths otherwise.  Columns
 * q >= num_sources stay HC_INF_DIST.  stats->ms: HIP events around the
 * levels, the fill of the matrix included. */
int hc_bfs_multi(hc_graph_t g, const int32_t *sources, int32_t num_sources,
                 int32_t *depth, int32_t max_epochs, hc_run_stats_t *stats,
                 hc_stream_t stream);

/* Copies column `column` of a batched result (num_nodes x 64, vertex-major)
 * into out (device, num_nodes int32), HC_INF_DIST mapped to -1 when
 * unreached_minus1 is set (BFS depths).  Asynchronous on stream. */
int hc_multi_extract(const int32_t *matrix, int32_t num_nodes, int32_t column,
                     int32_t unreached_minus1, int32_t *out, hc_stream_t stream);

/* Order-independent 64-bit checksum of a result (no counterpart in the
 * reference; for checking results against a run of the same query alone
 * without copying them to the host): the wrapping sum over v < num_nodes of
 * mix(v, data[v * stride + column]), stride 1 (a single-source array) or
 * 64 (a column of a batched matrix), HC_INF_DIST read as -1 when
 * unreached_minus1 is set; mix is splitmix64's finaliser of
 * (v << 32) ^ (uint32_t)value.  Waits for the result. */
int hc_checksum_i32(const int32_t *data, int32_t num_nodes, int32_t stride, int32_t column,
                    int32_t unreached_minus1, uint64_t *sum, hc_stream_t stream);

/* hc_checksum_i32 of columns 0 .. num_columns-1 of a batched [num_nodes][64]
 * result, all in one pass over the matrix (sums[c] for column c).  Waits
 * for the sums.  No counterpart in the reference. */
int hc_checksum_multi(const int32_t *matrix, int32_t num_nodes, int32_t num_columns,
                      int32_t unreached_minus1, uint64_t *sums, hc_stream_t stream);

/* Many graphs resident in HBM as one batch: concatenated 4-byte offsets and
 * neighbour ids (each graph below 2^32 edges), descriptors and a
 * longest-first work order.  Graph i's vertex v is batch vertex
 * node_offsets[i] + v. */
typedef struct hc_graph_batch_s *hc_graph_batch_t;
int hc_graph_batch_upload(const hc_host_graph_t *graphs, int32_t num_graphs,
                          hc_graph_batch_t *out);
int hc_graph_batch_free(hc_graph_batch_t b);
int hc_graph_batch_info(hc_graph_batch_t b, int32_t *num_graphs,
                        int64_t *total_nodes, int64_t *total_edges,
                        size_t *device_bytes);
/* out: num_graphs + 1 prefix sums of the vertex counts. */
int hc_graph_batch_node_offsets(hc_graph_batch_t b, int64_t *out);

/* One direction-optimizing BFS per graph (bfs_do_cpu's rule with alpha,
 * beta; 0 selects 15 and 18), sources[i] (host) in graph i, all graphs in
 * one persistent launch: each workgroup runs whole BFSs with the bitmaps of
 * its graph in LDS.  depth (and parents, may be NULL): device arrays of
 * total_nodes int32 laid out by node_offsets.  Depths are identical to
 * hc_bfs per graph; parents are valid BFS parents.  stats: push_epochs /
 * pull_epochs summed over graphs, ms of the launch. */
int hc_bfs_many(hc_graph_batch_t b, const int32_t *sources, int32_t *depth,
                int32_t *parents, int32_t alpha, int32_t beta,
                hc_run_stats_t *stats, hc_stream_t stream);

/* hc_bfs_many without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as hc_bfs_many; its
 * scratch (work counter, sources, bitmaps of graphs above 2^17 vertices) is
 * a fixed 32.25 MiB array per GPU that the library holds until the process
 * ends.  Limits: at most 65536 graphs, and a largest graph whose three
 * bitmaps (12 bytes per 32 vertices) fit in 32 MiB (about 89 M vertices);
 * above them the call fails with HC_ERR_INVALID.  Fewer workgroups run when
 * the bitmaps of a full grid do not fit.  Calls that overlap on one GPU
 * share the scratch and race: the caller runs one at a time per GPU.
 * stats->ms is the host-observed time of the launch. */
int hc_bfs_many_noalloc(hc_graph_batch_t b, const int32_t *sources,
                        int32_t *depth, int32_t *parents, int32_t alpha,
                        int32_t beta, hc_run_stats_t *stats,
                        hc_stream_t stream);


#ifdef __cplusplus
}
#endif

#endif /* HC_BATCH_H */
