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

#ifdef __cplusplus
}
#endif

#endif /* HC_BATCH_H */
