/*
 * oracle.h -- CPU restatement of the reference's algorithms on the graph
 * epoch path.  TEST INFRASTRUCTURE ONLY: linked by tests/, the smoke check and
 * the CPU-baseline leg of bench.py, never by the library under test.
 *
 * Graph arguments are a CSR: n vertices, index[n+1] (int64), nbr[index[n]]
 * (int32) and, for SSSP, wt[index[n]] (int32 >= 0).  Distances are int32 with
 * the saturating sentinel OR_INF (the reference's float infinity,
 * src/graph.cuh:32, has no int32 counterpart).
 */
#ifndef HC_ORACLE_H
#define HC_ORACLE_H

#include <stdint.h>

#define OR_INF 0x3FFFFFFF

#ifdef __cplusplus
extern "C" {
#endif

int32_t or_version(void);

/* SSSP */
void or_sssp_pull_serial(int32_t n, const int64_t *index, const int32_t *nbr,
                         const int32_t *wt, int32_t src, int32_t *dist);
void or_sssp_dijkstra(int32_t n, const int64_t *index, const int32_t *nbr,
                      const int32_t *wt, int32_t src, int32_t *dist);
int32_t or_sssp_jacobi_epoch(int32_t n, const int64_t *index,
                             const int32_t *nbr, const int32_t *wt,
                             int32_t start, int32_t end, const int32_t *din,
                             int32_t *dout);
double or_sssp_pull_omp(int32_t n, const int64_t *index, const int32_t *nbr,
                        const int32_t *wt, const int32_t *init, int32_t *dist,
                        int32_t *epochs, int32_t nthreads);

/* BFS */
void or_bfs_depth(int32_t n, const int64_t *index, const int32_t *nbr,
                  int32_t src, int32_t *depth);
int64_t or_verify_parents(int32_t n, const int32_t *depth,
                          const int32_t *parents, int32_t src);
int32_t or_bfs_pull_epoch(int32_t n, const int64_t *index, const int32_t *nbr,
                          int32_t *parents, int32_t start, int32_t end,
                          const uint32_t *frontier, uint32_t *next);
double or_bfs_do(int32_t n, const int64_t *index, const int32_t *nbr,
                 int32_t src, int32_t alpha, int32_t beta, int32_t *parents,
                 int8_t *dirs, int32_t max_dirs, int32_t *ndirs,
                 int32_t nthreads);

/* Certificate of k distance/depth vectors (column q of D, D[u*ld+q]) from
 * src[q]: bad[q] = vertices breaking it; returns the total, -1 when a weight
 * is < 1 or a source is out of range.  wt NULL = unit weights (BFS). */
int64_t or_certify_paths(int32_t n, const int64_t *index, const int32_t *nbr,
                         const int32_t *wt, int32_t k, const int32_t *src,
                         const int32_t *D, int64_t ld, int32_t unreached,
                         int64_t *bad);

/* PageRank */
void or_pr_jacobi_f32(int32_t n, const int64_t *index, const int32_t *nbr,
                      float damping, int32_t iters, float *scores);
void or_pr_jacobi_f64(int32_t n, const int64_t *index, const int32_t *nbr,
                      double damping, int32_t iters, double *scores);
int32_t or_pr_gauss_seidel_f32(int32_t n, const int64_t *index,
                               const int32_t *nbr, int32_t max_iters,
                               float *scores);
int32_t or_pr_jacobi_rule_f32(int32_t n, const int64_t *index, const int32_t *nbr,
                              float damping, int32_t max_iters, float *scores);
int32_t or_pr_jacobi_rule_f32_order(int32_t n, const int64_t *index, const int32_t *nbr,
                                    float damping, int32_t max_iters, int32_t reverse,
                                    float *scores);

/* Segmentation */
void or_equal_edge_ranges(int32_t n, const int64_t *index, int32_t segments,
                          int32_t chunk, int32_t *out);

#ifdef __cplusplus
}
#endif

#endif
