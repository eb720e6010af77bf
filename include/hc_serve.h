/*
 * hc_serve.h -- serving loop: graphs resident in HBM, BFS / SSSP queries
 * admitted by device-memory footprint and batched as they arrive.
 *
 * No reference counterpart (the reference runs one query per driver call and
 * uploads the graph each time, src/kernels/gpu/sssp.cuh:42-100).  A server
 * owns `slots` worker threads, each with its own HIP stream and its own
 * handle on every graph (hc_graph_share), so queries on different slots
 * overlap on the device.  A free slot takes the oldest waiting query and
 * every other waiting query of the same graph and kind, up to max_batch:
 * one query runs the single-source path (hc_bfs / hc_sssp), two or more run
 * one batched sweep (hc_bfs_multi / hc_sssp_multi).  Before it runs, a
 * batch reserves its result memory from the budget; when the budget is
 * spent the slot trims its buffer pool, evicts idle dynamic graphs, moves
 * finished but unreleased results to host memory, and otherwise waits for
 * running sweeps.  Admission is in arrival order.  Results are
 * bit-identical to running each query alone.
 */
#ifndef HC_SERVE_H
#define HC_SERVE_H

#include "hc_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

enum hc_query_kind { HC_QUERY_BFS = 0, HC_QUERY_SSSP = 1 };

typedef struct hc_server_opts {
    int32_t slots;         /* worker threads / streams (default 4)            */
    int32_t max_batch;     /* queries per batched sweep, 1..64 (default 64)   */
    int64_t memory_budget; /* HBM bytes for results and dynamic graphs
                              (default: half of the free HBM)                 */
    int32_t min_batch_sssp; /* a batched SSSP sweep runs when at least this
                               many SSSP queries of one graph wait; fewer run
                               one at a time.  0 (default): from the graph's
                               measured costs (see calibrate), else 8          */
    int32_t min_batch_bfs;  /* the same for BFS (0: measured, else 64)         */
    int32_t max_wait_ms;    /* queries of one graph and kind arriving fast
                               enough to gather 2 x its min_batch of them (at
                               most max_batch) within this many ms of the
                               oldest one's arrival wait (while other queries
                               run) until a full batch is there or the oldest
                               has waited this long, then run as one sweep;
                               slower classes run at once (default 500; < 0:
                               never wait)                                     */
    int32_t calibrate;      /* 0 (default): when a resident graph is added,
                               time one single-source query and one 64-wide
                               batch of each kind on it; a class then batches
                               from ceil(batch ms / single ms) waiting queries
                               (never, when that exceeds max_batch).  < 0: no
                               timing, the fixed defaults above                */
    int32_t checksum;       /* 1: after each sweep, before its slot takes more
                               work, the slot computes the hc_checksum_i32 of
                               every result of the sweep on the device (one
                               pass over a batched result) and keeps it for
                               hc_server_checksum; the sweep's latency_ms
                               ends before this pass (default 0)               */
} hc_server_opts_t;

typedef struct hc_query_info {
    int32_t status;       /* HC_OK, or the error of the run                   */
    int32_t kind;
    int32_t graph;
    int32_t source;
    int32_t batch_size;   /* queries in the sweep that answered this one      */
    int32_t slot;
    int64_t num_nodes;
    int64_t num_edges;    /* directed edges of the graph                       */
    double queue_ms;      /* submit -> start of its sweep                     */
    double run_ms;        /* sweep start -> result in HBM (host observed)     */
    double latency_ms;    /* submit -> result in HBM                          */
} hc_query_info_t;

typedef struct hc_server_stats {
    int64_t submitted, completed, failed;
    int64_t batches, batched_queries;
    int64_t uploads, evictions;
    int64_t spills;             /* finished results moved to host memory     */
    int64_t bytes_in_use, bytes_peak;
} hc_server_stats_t;

typedef struct hc_server_s *hc_server_t;

int hc_server_create(const hc_server_opts_t *opts, hc_server_t *out);
/* Waits for every submitted query, then frees everything. */
int hc_server_destroy(hc_server_t s);

/* Registers a graph.  resident = 1: uploaded now and kept.  resident = 0:
 * uploaded when a query first needs it, evicted when idle and memory is
 * needed; the caller keeps g's arrays alive until hc_server_destroy. */
int hc_server_add_graph(hc_server_t s, const hc_host_graph_t *g, int32_t resident,
                        int32_t *graph_id);

/* Device ms of one single-source query and of one 64-wide batched sweep of
 * each kind (index HC_QUERY_BFS / HC_QUERY_SSSP) measured when graph_id was
 * added; 0 where not measured (dynamic graphs, calibrate < 0, unweighted
 * SSSP). */
int hc_server_graph_costs(hc_server_t s, int32_t graph_id, double *single_ms,
                          double *batch_ms);

/* Records knob key = value for queries on graph graph_id
 * (hc_graph_set_tuning on its device handle and the slots' shares of it;
 * dynamic graphs take it again at every upload), e.g. the autotuner's
 * cached choice for that graph.  No counterpart in the reference. */
int hc_server_set_graph_tuning(hc_server_t s, int32_t graph_id, const char *key,
                               int64_t value);

/* Bytes the budget charges for graph g: for a dynamic graph, its CSR plus
 * the workspaces of one handle per slot (hc_graph_share) and the copies the
 * graph builds on first use (column-blocked CSR), at their upper bounds;
 * for a resident graph (whose CSR is outside the budget), the workspaces
 * and copies only.  Resident graphs are charged when added. */
int hc_server_graph_bytes(hc_server_t s, const hc_host_graph_t *g, int32_t resident,
                          int64_t *bytes);

/* Enqueues a query; returns at once with a ticket. */
int hc_server_submit(hc_server_t s, int32_t graph_id, int32_t kind, int32_t source,
                     int64_t *ticket);
/* Blocks until the query has finished; fills info (may be NULL). */
int hc_server_wait(hc_server_t s, int64_t ticket, hc_query_info_t *info);
/* Copies the finished query's result to host memory: num_nodes int32 (BFS
 * depth, -1 unreached; SSSP distance, HC_INF_DIST unreachable). */
int hc_server_fetch(hc_server_t s, int64_t ticket, int32_t *host_out);
/* The hc_checksum_i32 checksum of the finished query's result, as
 * hc_server_fetch would return it (BFS unreached read as -1): the value the
 * slot computed after the sweep when opts.checksum is set, else computed
 * now where the result lies, without copying it to the host.  Waits for the
 * query.  No counterpart in the reference. */
int hc_server_checksum(hc_server_t s, int64_t ticket, uint64_t *sum);
/* Frees the query's result; the ticket becomes invalid. */
int hc_server_release(hc_server_t s, int64_t ticket);
int hc_server_get_stats(hc_server_t s, hc_server_stats_t *out);

#ifdef __cplusplus
}
#endif

#endif /* HC_SERVE_H */
