/*
 * oracle.c -- CPU restatement of chrisliu/hetero-compute's graph epoch
 * algorithms, used as the parity checker and CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function names the reference
 * file:line it restates.  Plain C11 + OpenMP; built with -ffp-contract=off so
 * float expressions round exactly as written.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int32_t or_version(void) { return 1; }

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* d + w saturated at OR_INF (d <= OR_INF, 0 <= w < 2^31). */
static inline int32_t relax(int32_t d, int32_t w) {
    uint32_t c = (uint32_t)d + (uint32_t)w;
    return (int32_t)(c < (uint32_t)OR_INF ? c : (uint32_t)OR_INF);
}

/* ------------------------------------------------------------------------ */
/* SSSP                                                                      */
/* ------------------------------------------------------------------------ */

/* sssp_pull_cpu_serial (src/kernels/cpu/sssp.cuh:111-137): in-place pull
 * sweeps in vertex order until a sweep changes nothing. */
void or_sssp_pull_serial(int32_t n, const int64_t *index, const int32_t *nbr,
                         const int32_t *wt, int32_t src, int32_t *dist) {
    for (int32_t i = 0; i < n; i++) dist[i] = OR_INF;
    dist[src] = 0;
    int64_t updates;
    do {
        updates = 0;
        for (int32_t v = 0; v < n; v++) {
            for (int64_t i = index[v]; i < index[v + 1]; i++) {
                int32_t prop = relax(dist[nbr[i]], wt[i]);
                if (prop < dist[v]) {
                    dist[v] = prop;
                    updates++;
                }
            }
        }
    } while (updates != 0);
}

/* sssp_dijkstras_cpu_serial (src/kernels/cpu/sssp.cuh:74-109): binary heap
 * of (distance, vertex) with lazy deletion. */
typedef struct {
    int32_t d;
    int32_t v;
} hent;

static int hless(hent a, hent b) { return a.d < b.d || (a.d == b.d && a.v < b.v); }

void or_sssp_dijkstra(int32_t n, const int64_t *index, const int32_t *nbr,
                      const int32_t *wt, int32_t src, int32_t *dist) {
    for (int32_t i = 0; i < n; i++) dist[i] = OR_INF;
    dist[src] = 0;
    int64_t cap = 1024, sz = 0;
    hent *h = (hent *)malloc(sizeof(hent) * cap);
    h[sz++] = (hent){0, src};
    while (sz > 0) {
        hent top = h[0];
        h[0] = h[--sz];
        for (int64_t i = 0;;) { /* sift down */
            int64_t l = 2 * i + 1, r = l + 1, m = i;
            if (l < sz && hless(h[l], h[m])) m = l;
            if (r < sz && hless(h[r], h[m])) m = r;
            if (m == i) break;
            hent t = h[i];
            h[i] = h[m];
            h[m] = t;
            i = m;
        }
        if (top.d != dist[top.v]) continue;
        for (int64_t i = index[top.v]; i < index[top.v + 1]; i++) {
            int32_t nd = relax(top.d, wt[i]);
            int32_t v = nbr[i];
            if (nd < dist[v]) {
                dist[v] = nd;
                if (sz == cap) {
                    cap *= 2;
                    h = (hent *)realloc(h, sizeof(hent) * cap);
                }
                int64_t j = sz++;
                h[j] = (hent){nd, v};
                while (j > 0) { /* sift up */
                    int64_t p = (j - 1) / 2;
                    if (!hless(h[j], h[p])) break;
                    hent t = h[j];
                    h[j] = h[p];
                    h[p] = t;
                    j = p;
                }
            }
        }
    }
    free(h);
}

/* One Jacobi pull epoch over [start, end): dout[u] = min(din[u],
 * min din[v] + w); vertices outside the range are copied.  An in-place
 * (Gauss-Seidel) epoch, as the reference runs it
 * (src/kernels/cpu/sssp.cuh:155-182), lands between this result and the
 * fixpoint.  Returns the number of vertices changed. */
int32_t or_sssp_jacobi_epoch(int32_t n, const int64_t *index,
                             const int32_t *nbr, const int32_t *wt,
                             int32_t start, int32_t end, const int32_t *din,
                             int32_t *dout) {
    int32_t changed = 0;
    memcpy(dout, din, sizeof(int32_t) * (size_t)n);
    for (int32_t u = start; u < end; u++) {
        int32_t m = din[u];
        for (int64_t i = index[u]; i < index[u + 1]; i++) {
            int32_t c = relax(din[nbr[i]], wt[i]);
            if (c < m) m = c;
        }
        if (m != din[u]) changed++;
        dout[u] = m;
    }
    return changed;
}

/* sssp_pull_cpu + epoch_sssp_pull_cpu_one_to_one
 * (src/kernels/cpu/sssp.cuh:33-65, 155-182): every epoch opens a parallel
 * region; thread t takes vertices t, t+T, ... in place; the update count is
 * merged with an atomic add; loop until an epoch changes nothing.  Returns
 * the epoch-loop time in ms (timer scope of :46-59). */
double or_sssp_pull_omp(int32_t n, const int64_t *index, const int32_t *nbr,
                        const int32_t *wt, const int32_t *init, int32_t *dist,
                        int32_t *epochs, int32_t nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads)
    for (int32_t i = 0; i < n; i++) dist[i] = init[i];
    int32_t updated = 1, ep = 0;
    double t0 = now_ms();
    while (updated != 0) {
        updated = 0;
#pragma omp parallel num_threads(nthreads)
        {
            const int tid = omp_get_thread_num(), nt = omp_get_num_threads();
            int32_t local = 0;
            for (int32_t u = tid; u < n; u += nt) {
                int32_t nd = dist[u];
                for (int64_t i = index[u]; i < index[u + 1]; i++) {
                    int32_t prop = relax(dist[nbr[i]], wt[i]);
                    nd = prop < nd ? prop : nd;
                }
                if (nd != dist[u]) {
                    dist[u] = nd;
                    local++;
                }
            }
#pragma omp atomic
            updated += local;
        }
        ep++;
    }
    double ms = now_ms() - t0;
    if (epochs) *epochs = ep;
    return ms;
}

/* ------------------------------------------------------------------------ */
/* BFS                                                                       */
/* ------------------------------------------------------------------------ */

/* compute_proper_depth (tests/bfs/bfs_verify.cu:211-233): FIFO BFS. */
void or_bfs_depth(int32_t n, const int64_t *index, const int32_t *nbr,
                  int32_t src, int32_t *depth) {
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; i++) depth[i] = -1;
    depth[src] = 0;
    int64_t head = 0, tail = 0;
    q[tail++] = src;
    while (head < tail) {
        int32_t u = q[head++];
        for (int64_t i = index[u]; i < index[u + 1]; i++) {
            int32_t v = nbr[i];
            if (depth[v] == -1) {
                depth[v] = depth[u] + 1;
                q[tail++] = v;
            }
        }
    }
    free(q);
}

/* verify_parents (tests/bfs/bfs_verify.cu:133-197): a parent array is valid
 * iff unreached <=> parent -1, parents[src] == src and every other reached
 * vertex sits one level below its parent.  Returns the number of violations
 * (out-of-range parents count as violations too). */
int64_t or_verify_parents(int32_t n, const int32_t *depth,
                          const int32_t *parents, int32_t src) {
    int64_t errors = 0;
    for (int32_t u = 0; u < n; u++) {
        int32_t p = parents[u];
        if (p == -1) {
            if (depth[u] != -1) errors++;
        } else if (p < 0 || p >= n) {
            errors++;
        } else if (u == src) {
            if (p != src) errors++;
        } else if (depth[u] != depth[p] + 1) {
            errors++;
        }
    }
    return errors;
}

static inline int get_bit(const uint32_t *bm, int32_t v) {
    return (bm[v >> 5] >> (v & 31)) & 1u;
}

/* epoch_bfs_pull_cpu (src/kernels/cpu/bfs.cuh:305-335), serial order:
 * unvisited u takes its first neighbour (CSR order) in the frontier. */
int32_t or_bfs_pull_epoch(int32_t n, const int64_t *index, const int32_t *nbr,
                          int32_t *parents, int32_t start, int32_t end,
                          const uint32_t *frontier, uint32_t *next) {
    (void)n;
    int32_t count = 0;
    for (int32_t u = start; u < end; u++) {
        if (parents[u] != -1) continue;
        for (int64_t i = index[u]; i < index[u + 1]; i++) {
            int32_t v = nbr[i];
            if (get_bit(frontier, v)) {
                parents[u] = v;
                next[u >> 5] |= 1u << (u & 31);
                count++;
                break;
            }
        }
    }
    return count;
}

/* bfs_do_cpu<alpha, beta> (src/kernels/cpu/bfs.cuh:48-114) with its OpenMP
 * epochs: push by node (:212-244, compare-and-swap on parents, per-thread
 * windows flushed with fetch-and-add, window.h:106-143), pull (:305-335,
 * dynamic 1024-vertex chunks), and the window <-> bitmap conversions
 * (:346-379).  The compare-and-swap result is tested for success (the
 * reference tests the returned old value for truth, which also admits
 * vertices whose winning parent id is non-zero).  dirs receives 0 (push) or
 * 1 (pull) per epoch.  Returns the epoch-loop time in ms. */
#define OR_LOCAL_WINDOW 16384

double or_bfs_do(int32_t n, const int64_t *index, const int32_t *nbr,
                 int32_t src, int32_t alpha, int32_t beta, int32_t *parents,
                 int8_t *dirs, int32_t max_dirs, int32_t *ndirs,
                 int32_t nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const int32_t words = (n + 31) / 32;
    int32_t *win = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t *front = (uint32_t *)calloc((size_t)(words > 0 ? words : 1), 4);
    uint32_t *next = (uint32_t *)calloc((size_t)(words > 0 ? words : 1), 4);
#pragma omp parallel for num_threads(nthreads)
    for (int32_t i = 0; i < n; i++) parents[i] = -1;
    parents[src] = src;
    int64_t w_start = 0, w_end = 0, w_next = 0;
    win[w_next++] = src;
    w_start = w_end;
    w_end = w_next;
    int64_t edges_to_check = index[n];
    int64_t num_edges = index[src + 1] - index[src];
    int32_t nd = 0;
    double t0 = now_ms();
    while (w_start != w_end) {
        if (num_edges > edges_to_check / alpha) {
            /* PULL phase */
            int64_t prev_nodes;
            int64_t num_nodes = w_end - w_start;
#pragma omp parallel for num_threads(nthreads)
            for (int64_t i = w_start; i < w_end; i++)
                __sync_fetch_and_or(&front[win[i] >> 5], 1u << (win[i] & 31));
            do {
                prev_nodes = num_nodes;
                num_nodes = 0;
                int64_t cnt = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : cnt)
                {
#pragma omp for nowait schedule(dynamic, 1024)
                    for (int32_t u = 0; u < n; u++) {
                        if (parents[u] != -1) continue;
                        for (int64_t i = index[u]; i < index[u + 1]; i++) {
                            int32_t v = nbr[i];
                            if (get_bit(front, v)) {
                                parents[u] = v;
                                next[u >> 5] |= 1u << (u & 31);
                                cnt++;
                                break;
                            }
                        }
                    }
                }
                num_nodes = cnt;
                uint32_t *t = front;
                front = next;
                next = t;
                memset(next, 0, sizeof(uint32_t) * (size_t)words);
                if (nd < max_dirs) dirs[nd] = 1;
                nd++;
            } while (num_nodes > prev_nodes || num_nodes > n / beta);
            num_edges = 1;
            /* conv_bitmap_to_window */
            w_start = w_end = w_next = 0;
#pragma omp parallel num_threads(nthreads)
            {
                int32_t *loc = (int32_t *)malloc(sizeof(int32_t) * OR_LOCAL_WINDOW);
                int64_t lc = 0;
#pragma omp for nowait
                for (int32_t u = 0; u < n; u++) {
                    if (get_bit(front, u)) {
                        if (lc == OR_LOCAL_WINDOW) {
                            int64_t at = __sync_fetch_and_add(&w_next, lc);
                            memcpy(win + at, loc, sizeof(int32_t) * lc);
                            lc = 0;
                        }
                        loc[lc++] = u;
                    }
                }
                int64_t at = __sync_fetch_and_add(&w_next, lc);
                memcpy(win + at, loc, sizeof(int32_t) * lc);
                free(loc);
            }
            w_start = w_end;
            w_end = w_next;
            memset(front, 0, sizeof(uint32_t) * (size_t)words);
        } else {
            /* PUSH epoch */
            edges_to_check -= num_edges;
            num_edges = 0;
            int64_t ne = 0;
            const int64_t fs = w_start, fe = w_end;
#pragma omp parallel num_threads(nthreads) reduction(+ : ne)
            {
                int32_t *loc = (int32_t *)malloc(sizeof(int32_t) * OR_LOCAL_WINDOW);
                int64_t lc = 0;
#pragma omp for nowait
                for (int64_t f = fs; f < fe; f++) {
                    int32_t u = win[f];
                    for (int64_t i = index[u]; i < index[u + 1]; i++) {
                        int32_t v = nbr[i];
                        if (parents[v] == -1 &&
                            __sync_val_compare_and_swap(&parents[v], -1, u) == -1) {
                            if (lc == OR_LOCAL_WINDOW) {
                                int64_t at = __sync_fetch_and_add(&w_next, lc);
                                memcpy(win + at, loc, sizeof(int32_t) * lc);
                                lc = 0;
                            }
                            loc[lc++] = v;
                            ne += index[v + 1] - index[v];
                        }
                    }
                }
                int64_t at = __sync_fetch_and_add(&w_next, lc);
                memcpy(win + at, loc, sizeof(int32_t) * lc);
                free(loc);
            }
            num_edges = ne;
            w_start = w_end;
            w_end = w_next;
            if (nd < max_dirs) dirs[nd] = 0;
            nd++;
        }
    }
    double ms = now_ms() - t0;
    if (ndirs) *ndirs = nd;
    free(win);
    free(front);
    free(next);
    return ms;
}

/* ------------------------------------------------------------------------ */
/* Shortest-path certificate                                                 */
/* ------------------------------------------------------------------------ */

/* Checks k candidate distance vectors at once, in O(m k) with no search:
 * column q of D (D[u * ld + q], u < n) claims the distances from src[q]
 * under the pull relaxation d[u] = min over row u of d[v] + w (the rule of
 * epoch_sssp_pull_cpu_one_to_one, src/kernels/cpu/sssp.cuh:155-182, and of
 * the pull BFS epochs, src/kernels/cpu/bfs.cuh:305-335, with w = 1).  A
 * column is accepted iff
 *   (a) every entry is `unreached` or lies in [0, OR_INF), and d[src] = 0;
 *   (b) no edge is violated: d[u] <= d[v] + w for every (v, w) of row u
 *       (unreached counts as OR_INF, sums saturate at OR_INF);
 *   (c) every reached u != src has a tight edge: some (v, w) of row u with
 *       d[v] + w == d[u].
 * With every weight >= 1 these are exactly the true distances: (b) makes d
 * a feasible potential, so d <= the true distance; (c) gives each reached
 * vertex a predecessor with a strictly smaller value, a chain that can only
 * end at src, i.e. a path of length d[u].  Unreached vertices then have no
 * reached neighbour (by (b)), so the reached set is src's component.
 * wt NULL: unit weights (BFS depths).  bad[q] receives the number of
 * vertices whose entry in column q breaks (a)-(c).  Returns the total, or
 * -1 when a weight is below 1 or a source lies outside [0, n) (the proof
 * above needs positive weights; graphs with zero weights are checked
 * against Dijkstra instead). */
int64_t or_certify_paths(int32_t n, const int64_t *index, const int32_t *nbr,
                         const int32_t *wt, int32_t k, const int32_t *src,
                         const int32_t *D, int64_t ld, int32_t unreached,
                         int64_t *bad) {
    for (int32_t q = 0; q < k; q++) {
        bad[q] = 0;
        if (src[q] < 0 || src[q] >= n) return -1;
    }
    const int64_t m = index[n];
    if (wt) {
        int64_t low = 0;
#pragma omp parallel for reduction(+ : low)
        for (int64_t i = 0; i < m; i++) low += wt[i] < 1;
        if (low) return -1;
    }
    int64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        int32_t *du = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
        uint8_t *ok = (uint8_t *)malloc((size_t)k);   /* (a) holds, no violation yet */
        uint8_t *tight = (uint8_t *)malloc((size_t)k);
        int64_t *mine = (int64_t *)calloc((size_t)k, sizeof(int64_t));
#pragma omp for schedule(dynamic, 512)
        for (int32_t u = 0; u < n; u++) {
            const int32_t *row = D + (int64_t)u * ld;
            for (int32_t q = 0; q < k; q++) {
                const int32_t x = row[q];
                ok[q] = 1;
                tight[q] = 0;
                if (x == unreached) {
                    du[q] = OR_INF;
                } else if (x < 0 || x >= OR_INF) {
                    ok[q] = 0;
                    du[q] = OR_INF;
                } else {
                    du[q] = x;
                }
                if (u == src[q] && du[q] != 0) ok[q] = 0;
            }
            for (int64_t i = index[u]; i < index[u + 1]; i++) {
                const int32_t *rv = D + (int64_t)nbr[i] * ld;
                const int32_t w = wt ? wt[i] : 1;
                for (int32_t q = 0; q < k; q++) {
                    int32_t dv = rv[q];
                    if (dv == unreached || dv < 0 || dv >= OR_INF) dv = OR_INF;
                    const int32_t c = relax(dv, w);
                    if (c < du[q]) ok[q] = 0;
                    if (c == du[q]) tight[q] = 1;
                }
            }
            for (int32_t q = 0; q < k; q++) {
                const int reached_other = du[q] < OR_INF && u != src[q];
                if (!ok[q] || (reached_other && !tight[q])) mine[q]++;
            }
        }
        for (int32_t q = 0; q < k; q++) {
            if (mine[q]) {
#pragma omp atomic
                bad[q] += mine[q];
            }
            total += mine[q];
        }
        free(du);
        free(ok);
        free(tight);
        free(mine);
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* PageRank                                                                  */
/* ------------------------------------------------------------------------ */

/* Jacobi form of the per-vertex formula of epoch_pr_pull_cpu_one_to_one
 * (src/kernels/cpu/pr.cuh:165-175) in FP32: score'[u] = (1-d)/N +
 * d * sum (score[v] / deg(v)), the sum sequential in CSR order, start 1/N
 * (:39-42).  Runs exactly `iters` iterations. */
void or_pr_jacobi_f32(int32_t n, const int64_t *index, const int32_t *nbr,
                      float damping, int32_t iters, float *scores) {
    float *cur = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    float *nxt = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    const float init = 1.0f / (float)n;
    const float base = (1.0f - damping) / (float)n;
    for (int32_t i = 0; i < n; i++) cur[i] = init;
    for (int32_t it = 0; it < iters; it++) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t u = 0; u < n; u++) {
            float total = 0.0f;
            for (int64_t i = index[u]; i < index[u + 1]; i++) {
                int32_t v = nbr[i];
                total += cur[v] / (float)(index[v + 1] - index[v]);
            }
            nxt[u] = base + damping * total;
        }
        float *t = cur;
        cur = nxt;
        nxt = t;
    }
    memcpy(scores, cur, sizeof(float) * (size_t)n);
    free(cur);
    free(nxt);
}

/* The same Jacobi iteration in FP64 (the accuracy reference). */
void or_pr_jacobi_f64(int32_t n, const int64_t *index, const int32_t *nbr,
                      double damping, int32_t iters, double *scores) {
    double *cur = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double *nxt = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    const double base = (1.0 - damping) / (double)n;
    for (int32_t i = 0; i < n; i++) cur[i] = 1.0 / (double)n;
    for (int32_t it = 0; it < iters; it++) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int32_t u = 0; u < n; u++) {
            double total = 0.0;
            for (int64_t i = index[u]; i < index[u + 1]; i++) {
                int32_t v = nbr[i];
                total += cur[v] / (double)(index[v + 1] - index[v]);
            }
            nxt[u] = base + damping * total;
        }
        double *t = cur;
        cur = nxt;
        nxt = t;
    }
    memcpy(scores, cur, sizeof(double) * (size_t)n);
    free(cur);
    free(nxt);
}

/* pr_pull_cpu as written (src/kernels/cpu/pr.cuh:33-65, 157-187), serial
 * vertex order: in-place updates, a vertex counts as updated when its score
 * moves by more than 5e-8 (std::fabs), loop until no vertex counts or
 * max_iters sweeps.  Returns the sweeps run. */
int32_t or_pr_gauss_seidel_f32(int32_t n, const int64_t *index,
                               const int32_t *nbr, int32_t max_iters,
                               float *scores) {
    const float kDamp = 0.85f, eps = 0.00000005f;
    const float base = (1.0f - kDamp) / (float)n;
    for (int32_t i = 0; i < n; i++) scores[i] = 1.0f / (float)n;
    int32_t sweeps = 0, updated = 1;
    while (updated != 0 && sweeps < max_iters) {
        updated = 0;
        for (int32_t u = 0; u < n; u++) {
            float total = 0.0f;
            for (int64_t i = index[u]; i < index[u + 1]; i++) {
                int32_t v = nbr[i];
                total += scores[v] / (float)(index[v + 1] - index[v]);
            }
            float ns = base + kDamp * total;
            if (fabsf(ns - scores[u]) > eps) {
                scores[u] = ns;
                updated++;
            }
        }
        sweeps++;
    }
    return sweeps;
}

/* The stopping rule of pr_pull_cpu as written (src/kernels/cpu/pr.cuh:33-65,
 * 165-181) in Jacobi form: every vertex computes its new score from the
 * previous sweep's scores; it takes the new score only when that moves it
 * by more than 5e-8 (std::fabs in FP32), and the loop stops after a sweep in
 * which no vertex moved, or after max_iters sweeps.  Returns the sweeps run. */
int32_t or_pr_jacobi_rule_f32(int32_t n, const int64_t *index, const int32_t *nbr,
                              float damping, int32_t max_iters, float *scores) {
    return or_pr_jacobi_rule_f32_order(n, index, nbr, damping, max_iters, 0, scores);
}

/* The same with each vertex's FP32 sum taken over its neighbours in reverse
 * CSR order when `reverse` is set: the rule's 5e-8 test is decided on FP32
 * values, so another summation order (the device's compensated sums are
 * one) moves some vertices across it. */
int32_t or_pr_jacobi_rule_f32_order(int32_t n, const int64_t *index, const int32_t *nbr,
                                    float damping, int32_t max_iters, int32_t reverse,
                                    float *scores) {
    const float eps = 0.00000005f;
    float *cur = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    float *nxt = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    const float base = (1.0f - damping) / (float)n;
    for (int32_t i = 0; i < n; i++) cur[i] = 1.0f / (float)n;
    int32_t sweeps = 0;
    long long updated = 1;
    while (updated != 0 && sweeps < max_iters) {
        updated = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : updated)
        for (int32_t u = 0; u < n; u++) {
            float total = 0.0f;
            const int64_t b = index[u], e = index[u + 1];
            for (int64_t k = 0; k < e - b; k++) {
                const int32_t v = nbr[reverse ? e - 1 - k : b + k];
                total += cur[v] / (float)(index[v + 1] - index[v]);
            }
            const float ns = base + damping * total;
            const int moved = fabsf(ns - cur[u]) > eps;
            nxt[u] = moved ? ns : cur[u];
            updated += moved;
        }
        float *t = cur;
        cur = nxt;
        nxt = t;
        sweeps++;
    }
    memcpy(scores, cur, sizeof(float) * (size_t)n);
    free(cur);
    free(nxt);
    return sweeps;
}

/* ------------------------------------------------------------------------ */
/* Segmentation                                                              */
/* ------------------------------------------------------------------------ */

/* compute_equal_edge_ranges (src/graph.cuh:156-201) as written, except that
 * no boundary is written past out[segments] and boundaries the reference
 * leaves unset stay -1. */
void or_equal_edge_ranges(int32_t n, const int64_t *index, int32_t segments,
                          int32_t chunk, int32_t *out) {
    for (int32_t i = 0; i <= segments; i++) out[i] = -1;
    out[0] = 0;
    int64_t avg = index[n] / segments;
    int32_t end_id = 0, seg_id = 0;
    int64_t seg_deg = 0;
    while (end_id != n) {
        seg_deg += index[end_id + 1] - index[end_id];
        if (seg_deg >= avg) {
            if (seg_id + 1 <= segments) out[seg_id + 1] = end_id;
            seg_deg = 0;
            seg_id++;
        }
        end_id++;
    }
    if (seg_id != segments && seg_id + 1 <= segments) out[seg_id + 1] = end_id;
    for (int32_t i = 1; i < segments; i++) {
        if (out[i] < 0) continue;
        int32_t diff = out[i] % chunk;
        if (diff != 0) {
            if (diff < chunk / 2)
                out[i] = out[i] / chunk * chunk;
            else
                out[i] = (out[i] / chunk + 1) * chunk;
        }
    }
}
