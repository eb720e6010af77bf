// PageRank pull epochs for gfx950 (Jacobi form).
//
// score'[u] = (1-d)/N + d * sum_{v in N(u)} contrib[v], contrib[v] =
// score[v]/deg(v): the per-vertex formula of epoch_pr_pull_cpu_one_to_one
// (src/kernels/cpu/pr.cuh:165-175) with
//   * the neighbour sum in FP32 with Kahan compensation per lane and
//     TwoSum-combined lane and chunk partials (fixed order: deterministic);
//   * contrib computed once per vertex as (float)((double)score * inv_deg)
//     with inv_deg = 1/deg in FP64, fused into the epoch that writes score;
//   * the L1 residual sum |score' - score| in FP64: one partial per
//     workgroup, summed in workgroup order by a one-workgroup controller
//     kernel, which also finishes the chunked hub vertices and takes the
//     stopping decision (residual < tol or max_iters reached);
//   * or, with rule 1, the stopping rule of pr_pull_cpu as written
//     (src/kernels/cpu/pr.cuh:165-181) in Jacobi form: a vertex takes its
//     new score only when it moves by more than 5e-8 (FP32 fabs), and the
//     loop stops after an iteration in which no vertex moved.
// Scores start at 1/N, as pr_pull_cpu does (src/kernels/cpu/pr.cuh:38-42).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "frontier.h"

namespace hc {

struct KSum {
    float hi, lo;  // value ~ hi + lo
};

__device__ __forceinline__ KSum two_sum_combine(KSum a, KSum b) {
    float s = a.hi + b.hi;
    float bp = s - a.hi;
    float err = (a.hi - (s - bp)) + (b.hi - bp);
    return KSum{s, a.lo + b.lo + err};
}

template <int G>
__device__ __forceinline__ KSum group_ksum(KSum k) {
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) {
        KSum o{__shfl_xor(k.hi, s, G), __shfl_xor(k.lo, s, G)};
        // Combine in a lane-order-independent way: the lower lane's value is
        // the left operand on both sides of the exchange.
        bool low = (lane_id() & s) == 0;
        k = low ? two_sum_combine(k, o) : two_sum_combine(o, k);
    }
    return k;
}

// Kahan-compensated sum of contrib over edges [b, e) strided by G, in edge
// order per lane.  Trips of four edges with predicated loads, and the next
// trip's ids loaded before this trip's gathers (one memory round trip per
// trip); padding lanes skip the update, so the sum is the same sequence of
// Kahan steps as a plain loop over the lane's edges.
template <int G, typename OffT>
__device__ __forceinline__ KSum row_ksum(const GraphView<OffT> &g,
                                         const Staged<float> &contrib,
                                         OffT b, OffT e, int lane) {
    float sum = 0.0f, comp = 0.0f;
    OffT i = b + OffT(lane);
    if (i >= e) return KSum{0.0f, 0.0f};
    int32_t v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const OffT k = i + OffT(r * G);
        v[r] = k < e ? ldnt(&g.adj[k]) : -1;
    }
    for (;;) {
        const OffT ni = i + OffT(4 * G);
        int32_t nv[4] = {-1, -1, -1, -1};
        if (ni < e) {  // short rows skip the prefetch instructions entirely
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const OffT k = ni + OffT(r * G);
                nv[r] = k < e ? ldnt(&g.adj[k]) : -1;
            }
        }
        float x[4];
#pragma unroll
        for (int r = 0; r < 4; r++) x[r] = v[r] >= 0 ? contrib[v[r]] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (v[r] < 0) continue;
            float y = x[r] - comp;
            float t = sum + y;
            comp = (t - sum) - y;
            sum = t;
        }
        if (ni >= e) break;
        for (int r = 0; r < 4; r++) v[r] = nv[r];
        i = ni;
    }
    return KSum{sum, -comp};
}

__device__ __forceinline__ KSum block_ksum(KSum k) {
    __shared__ float s_hi[16], s_lo[16];
    k = group_ksum<64>(k);
    __syncthreads();
    if (lane_id() == 0) {
        s_hi[threadIdx.x / kWave] = k.hi;
        s_lo[threadIdx.x / kWave] = k.lo;
    }
    __syncthreads();
    KSum r{s_hi[0], s_lo[0]};
    const int nw = blockDim.x / kWave;
    for (int i = 1; i < nw; i++) r = two_sum_combine(r, KSum{s_hi[i], s_lo[i]});
    return r;
}

__device__ __forceinline__ double block_sum_f64(double v) {
    __shared__ double s_part[16];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    __syncthreads();
    if (lane_id() == 0) s_part[threadIdx.x / kWave] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / kWave); i++) t += s_part[i];
    return t;
}

struct PrBufs {
    const double *inv_deg;
    float *hub_hi;         // per chunk partials
    float *hub_lo;
    uint32_t *hub_cnt;     // per hub: chunks finished this epoch
    double *hub_resid;     // per hub: |score' - score|
    double *partials;      // per workgroup residual partial
};

struct PrCtl {
    int32_t rule;       // 0: FP64 L1 residual < tol; 1: no vertex moved > 5e-8
    int32_t done;
    int32_t iter;
    int32_t max_iters;
    double tol;
    double residual;
    double alg_bytes;
    double epoch_bytes;
    long long t_start;  // device wall clock at init
    long long t_end;    // and at the end of the last iteration
};

struct PrEpoch {
    const float *s_in;
    const float *c_in;
    float *s_out;
    float *c_out;
    float base;
    float damp;
    int rule;  // PrCtl::rule
};

// Rule 0 returns |score' - score| (FP64); rule 1 keeps the old score unless
// the new one moves it by more than 5e-8 (epoch_pr_pull_cpu_one_to_one,
// src/kernels/cpu/pr.cuh:170-181) and returns 1 when it moved.
__device__ __forceinline__ double pr_finalize(const PrEpoch &ep, const double *inv_deg,
                                              int32_t u, KSum k) {
    float sum = k.hi + k.lo;
    float s_new = __fadd_rn(ep.base, __fmul_rn(ep.damp, sum));
    const float s_old = ep.s_in[u];
    if (ep.rule) {
        const bool moved = fabsf(__fsub_rn(s_new, s_old)) > 0.00000005f;
        const float s = moved ? s_new : s_old;
        ep.s_out[u] = s;
        ep.c_out[u] = (float)((double)s * inv_deg[u]);
        return moved ? 1.0 : 0.0;
    }
    ep.s_out[u] = s_new;
    ep.c_out[u] = (float)((double)s_new * inv_deg[u]);
    return fabs((double)s_new - (double)s_old);
}

template <int G, typename OffT>
__device__ __forceinline__ void pr_group(const PlanDev &p, int c, int k,
                                         const GraphView<OffT> &g,
                                         const Staged<float> &cr,
                                         const PrEpoch &ep, const PrBufs &b,
                                         double &racc) {
    constexpr int PER = kBlock / G;
    const int gid = threadIdx.x / G, lane = threadIdx.x % G;
    const int idx = k * PER + gid;
    if (idx < p.cnt[c]) {
        int32_t u = p.list[c] ? p.list[c][idx] : p.beg[c] + idx;
        KSum s = row_ksum<G>(g, cr, g.rbeg(u), g.rend(u), lane);
        if (G > 1) s = group_ksum<G>(s);
        if (lane == 0) racc += pr_finalize(ep, b.inv_deg, u, s);
    }
}

// One Jacobi sweep over the plan units and the zero-degree vertices.
// Returns this thread's residual contribution (hub residuals go to
// b.hub_resid instead, so the total does not depend on which workgroup
// finishes a hub).
// contrib is read-only within a sweep, so the LDS snapshot of its hub entries
// (staged from ep.c_in) is exact.
template <typename OffT>
__device__ double pr_sweep(const PlanDev &p, const GraphView<OffT> &g,
                           const PrEpoch &ep, const PrBufs &b, int32_t nstage,
                           float *lds) {
    const Staged<float> cr = stage(ep.c_in, nstage, lds);
    double racc = 0.0;
    const int gstride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < p.zero_cnt; i += gstride) {
        int32_t u = p.zero_list ? p.zero_list[i] : p.zero_beg + i;
        racc += pr_finalize(ep, b.inv_deg, u, KSum{0.0f, 0.0f});
    }
    const int total = p.unit_start[kNumClasses];
    for (int unit = blockIdx.x; unit < total; unit += gridDim.x) {
        int c = 0;
        while (unit >= p.unit_start[c + 1]) c++;
        const int k = unit - p.unit_start[c];
        if (c == 0) {
            HubChunk h = p.chunks[k];
            KSum s = block_ksum(row_ksum<kBlock>(g, cr, OffT(h.b), OffT(h.e),
                                                 threadIdx.x));
            // Chunk partials are combined by the controller kernel.
            if (threadIdx.x == 0) {
                b.hub_hi[k] = s.hi;
                b.hub_lo[k] = s.lo;
            }
        } else if (c == 1) {
            int32_t u = p.list[1] ? p.list[1][k] : p.beg[1] + k;
            KSum s = block_ksum(row_ksum<kBlock>(g, cr, g.rbeg(u), g.rend(u),
                                                 threadIdx.x));
            if (threadIdx.x == 0) racc += pr_finalize(ep, b.inv_deg, u, s);
        } else if (c == 2) {
            pr_group<64>(p, c, k, g, cr, ep, b, racc);
        } else if (c == 3) {
            pr_group<16>(p, c, k, g, cr, ep, b, racc);
        } else if (c == 4) {
            pr_group<4>(p, c, k, g, cr, ep, b, racc);
        } else {
            pr_group<1>(p, c, k, g, cr, ep, b, racc);
        }
    }
    return racc;
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void pr_step_kernel(
    PlanDev p, GraphView<OffT> g, float *s0, float *s1, float *c0, float *c1,
    PrBufs b, PrCtl *ctl, float base, float damp, int32_t nstage) {
    extern __shared__ float s_stagef[];
    if (ctl->done) return;
    const int it = ctl->iter;
    const bool odd = it & 1;
    PrEpoch ep{odd ? s1 : s0, odd ? c1 : c0, odd ? s0 : s1, odd ? c0 : c1, base, damp, ctl->rule};
    double r = block_sum_f64(pr_sweep(p, g, ep, b, nstage, s_stagef));
    if (threadIdx.x == 0) b.partials[blockIdx.x] = r;
}

// Hub vertices: combine their chunk partials in chunk order and finish
// them; then the fixed-order residual total (workgroup partials, then hub
// residuals).  One workgroup of kBlock threads; returns the total in every
// thread.
__device__ double pr_hubs_and_residual(const PlanDev &p, const PrEpoch &ep,
                                       const PrBufs &b, int nparts) {
    for (int h = threadIdx.x; h < p.num_hubs; h += kBlock) {
        const int32_t c0 = p.hub_chunk_start[h], c1 = p.hub_chunk_start[h + 1];
        KSum tot{0.0f, 0.0f};
        for (int32_t j = c0; j < c1; j++) tot = two_sum_combine(tot, KSum{b.hub_hi[j], b.hub_lo[j]});
        b.hub_resid[h] = pr_finalize(ep, b.inv_deg, p.chunks[c0].u, tot);
    }
    __syncthreads();
    __shared__ double s_tot[kBlock];
    double t = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) t += b.partials[i];
    for (int i = threadIdx.x; i < p.num_hubs; i += kBlock) t += b.hub_resid[i];
    s_tot[threadIdx.x] = t;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) s_tot[threadIdx.x] += s_tot[threadIdx.x + s];
        __syncthreads();
    }
    double r = s_tot[0];
    __syncthreads();
    return r;
}

// Controller of the iteration loop (one workgroup).
__global__ __launch_bounds__(kBlock) void pr_ctl_kernel(
    PlanDev p, float *s0, float *s1, float *c0, float *c1, PrBufs b, PrCtl *ctl,
    float base, float damp, int nparts) {
    if (ctl->done) return;
    const int it = ctl->iter;
    const bool odd = it & 1;
    PrEpoch ep{odd ? s1 : s0, odd ? c1 : c0, odd ? s0 : s1, odd ? c0 : c1, base, damp, ctl->rule};
    const double r = pr_hubs_and_residual(p, ep, b, nparts);
    if (threadIdx.x == 0) {
        // Rule 1: r counts the vertices that moved (exact in FP64).
        const bool converged = ctl->rule ? r == 0.0 : r < ctl->tol;
        ctl->residual = r;
        ctl->iter = it + 1;
        ctl->alg_bytes += ctl->epoch_bytes;
        ctl->t_end = wall_clock64();
        if (converged || it + 1 >= ctl->max_iters) ctl->done = 1;
    }
}

template <typename OffT>
__global__ __launch_bounds__(kBlock) void pr_init_kernel(
    GraphView<OffT> g, float *s0, float *c0, double *inv_deg, uint32_t *hub_cnt,
    int32_t num_hubs, PrCtl *ctl, int max_iters, double tol, double epoch_bytes, int rule) {
    const int stride = gridDim.x * kBlock;
    const float init = 1.0f / (float)g.n;
    for (int u = blockIdx.x * kBlock + threadIdx.x; u < g.n; u += stride) {
        int64_t d = int64_t(g.rend(u) - g.rbeg(u));
        double inv = d > 0 ? 1.0 / (double)d : 0.0;
        inv_deg[u] = inv;
        s0[u] = init;
        c0[u] = (float)((double)init * inv);
    }
    for (int h = blockIdx.x * kBlock + threadIdx.x; h < num_hubs; h += stride) hub_cnt[h] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->done = max_iters <= 0 ? 1 : 0;
        ctl->t_start = wall_clock64();
        ctl->t_end = ctl->t_start;
        ctl->iter = 0;
        ctl->max_iters = max_iters;
        ctl->tol = tol;
        ctl->rule = rule;
        ctl->residual = 0.0;
        ctl->alg_bytes = 0.0;
        ctl->epoch_bytes = epoch_bytes;
    }
}

// Epoch API: one sweep over a plan; its finish kernel completes the hubs and
// adds the fixed-order residual to *residual_out.
template <typename OffT>
__global__ __launch_bounds__(kBlock) void pr_epoch_kernel(PlanDev p, GraphView<OffT> g,
                                                          PrEpoch ep, PrBufs b,
                                                          int32_t nstage) {
    extern __shared__ float s_stagef[];
    double r = block_sum_f64(pr_sweep(p, g, ep, b, nstage, s_stagef));
    if (threadIdx.x == 0) b.partials[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void pr_epoch_finish_kernel(PlanDev p, PrEpoch ep,
                                                                 PrBufs b, int nparts,
                                                                 double *residual_out) {
    const double r = pr_hubs_and_residual(p, ep, b, nparts);
    if (threadIdx.x == 0 && residual_out) *residual_out += r;
}

// ---------------------------------------------------------------------------
// Vertex-mapped epochs of the pr_pull_gpu driver (src/kernels/gpu/pr.cuh:
// 42-118; epochs epoch_pr_pull_gpu_{one_to_one,warp_red,block_red}, :140-
// 318): a thread, a wavefront or a workgroup per vertex over [0, n) with the
// caller's launch shape, the Jacobi step and stopping measure of the fused
// sweep (pr_finalize).  One residual partial per workgroup.
// ---------------------------------------------------------------------------

// Kahan sum of contrib over edges [b, e) taken `stride` apart from b + lane.
template <typename OffT>
__device__ __forceinline__ KSum row_ksum_strided(const GraphView<OffT> &g, const float *contrib,
                                                 OffT b, OffT e, int lane, int stride) {
    float sum = 0.0f, comp = 0.0f;
    for (OffT i = b + OffT(lane); i < e; i += OffT(stride)) {
        float y = contrib[g.adj[i]] - comp;
        float t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    return KSum{sum, -comp};
}

template <int MAP, typename OffT>
__global__ void pr_pull_mapped_kernel(GraphView<OffT> g, PrEpoch ep, const double *inv_deg,
                                      double *partials) {
    double racc = 0.0;
    if (MAP == HC_MAP_THREAD) {
        const int32_t stride = gridDim.x * blockDim.x;
        for (int32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < g.n; u += stride)
            racc += pr_finalize(ep, inv_deg, u,
                                row_ksum<1>(g, unstaged(ep.c_in), g.rbeg(u), g.rend(u), 0));
    } else if (MAP == HC_MAP_WAVE) {
        const int32_t waves = gridDim.x * (blockDim.x / kWave);
        for (int32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < g.n;
             u += waves) {
            KSum k = group_ksum<64>(
                row_ksum<64>(g, unstaged(ep.c_in), g.rbeg(u), g.rend(u), lane_id()));
            if (lane_id() == 0) racc += pr_finalize(ep, inv_deg, u, k);
        }
    } else {
        for (int32_t u = blockIdx.x; u < g.n; u += gridDim.x) {
            KSum k = block_ksum(row_ksum_strided(g, ep.c_in, g.rbeg(u), g.rend(u),
                                                 (int)threadIdx.x, (int)blockDim.x));
            if (threadIdx.x == 0) racc += pr_finalize(ep, inv_deg, u, k);
        }
    }
    const double r = block_sum_f64(racc);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// Initial state of a pr_pull_gpu run: scores from `init` (or 1/N), their
// contributions and the FP64 inverse degrees.
template <typename OffT>
__global__ void pr_pull_init_kernel(GraphView<OffT> g, const float *init, float *s0, float *c0,
                                    double *inv_deg) {
    const int32_t stride = gridDim.x * blockDim.x;
    const float uniform = 1.0f / (float)g.n;
    for (int32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < g.n; u += stride) {
        const int64_t d = int64_t(g.rend(u) - g.rbeg(u));
        const double inv = d > 0 ? 1.0 / (double)d : 0.0;
        const float x = init ? init[u] : uniform;
        inv_deg[u] = inv;
        s0[u] = x;
        c0[u] = (float)((double)x * inv);
    }
}

__global__ void pr_pull_zero_kernel(double *x) {
    if (threadIdx.x == 0) *x = 0.0;
}

// Sum of the workgroup partials in workgroup order (one workgroup).
__global__ __launch_bounds__(kBlock) void pr_pull_total_kernel(const double *partials,
                                                              int nparts, double *total) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) v += partials[i];
    const double t = block_sum_f64(v);
    if (threadIdx.x == 0) *total = t;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct PrWs {
    KernelTimer kt;  // step-kernel launches of the current run
    float *s[2] = {nullptr, nullptr};
    float *c[2] = {nullptr, nullptr};
    double *inv_deg = nullptr;
    float *hub_hi = nullptr, *hub_lo = nullptr;
    uint32_t *hub_cnt = nullptr;
    double *hub_resid = nullptr;
    double *partials = nullptr;
    PrCtl *ctl = nullptr;
    int32_t chunk_cap = 0, hub_cap = 0, part_cap = 0;
    void *dmem = nullptr;
};

void free_pr_ws(hc_graph_s *g) {
    if (!g->ws_pr) return;
    auto *ws = static_cast<PrWs *>(g->ws_pr);
    if (ws->dmem) hipFree(ws->dmem);
    delete ws;
    g->ws_pr = nullptr;
}

static int get_pr_ws(hc_graph_s *g, int32_t chunks, int32_t hubs, PrWs **out) {
    auto *ws = static_cast<PrWs *>(g->ws_pr);
    const int32_t parts = grid_blocks();
    if (ws && ws->chunk_cap >= chunks && ws->hub_cap >= hubs && ws->part_cap >= parts) {
        *out = ws;
        return HC_OK;
    }
    free_pr_ws(g);
    ws = new PrWs();
    const size_t n = std::max(g->n, 1);
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t sz_f = al(sizeof(float) * n), sz_d = al(sizeof(double) * n);
    size_t sz_ch = al(sizeof(float) * std::max(chunks, 1));
    size_t sz_h = al(sizeof(double) * std::max(hubs, 1));
    size_t sz_p = al(sizeof(double) * parts);
    size_t total = 4 * sz_f + sz_d + 2 * sz_ch + 2 * sz_h + sz_p + al(sizeof(PrCtl));
    char *base = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&base), total) != hipSuccess) {
        delete ws;
        return fail(HC_ERR_OOM, "PageRank workspace of %zu bytes failed", total);
    }
    size_t o = 0;
    for (int i = 0; i < 2; i++) {
        ws->s[i] = reinterpret_cast<float *>(base + o);
        o += sz_f;
        ws->c[i] = reinterpret_cast<float *>(base + o);
        o += sz_f;
    }
    ws->inv_deg = reinterpret_cast<double *>(base + o);
    o += sz_d;
    ws->hub_hi = reinterpret_cast<float *>(base + o);
    o += sz_ch;
    ws->hub_lo = reinterpret_cast<float *>(base + o);
    o += sz_ch;
    ws->hub_cnt = reinterpret_cast<uint32_t *>(base + o);
    o += sz_h;
    ws->hub_resid = reinterpret_cast<double *>(base + o);
    o += sz_h;
    ws->partials = reinterpret_cast<double *>(base + o);
    o += sz_p;
    ws->ctl = reinterpret_cast<PrCtl *>(base + o);
    ws->dmem = base;
    ws->chunk_cap = std::max(chunks, 1);
    ws->hub_cap = std::max(hubs, 1);
    ws->part_cap = parts;
    g->ws_pr = ws;
    *out = ws;
    return HC_OK;
}

static long long pr_epoch_bytes(const hc_graph_s *g, const PlanHost *ph) {
    const PlanDev &p = ph->dev;
    long long rows = p.num_hubs;
    for (int c = 1; c < kNumClasses; c++) rows += p.cnt[c];
    const long long ob = g->off64 ? 8 : 4;
    // offsets; id + gathered contrib per edge; score in/out, contrib out and
    // FP64 inverse degree per row; zero-degree rows: score in/out + contrib.
    return ob * (rows + 1) + 8ll * g->m + 20ll * rows + 12ll * p.zero_cnt;
}

template <typename OffT>
static int run_pr(hc_graph_s *g, float *scores, float damp, double tol,
                  int max_iters, int rule, hc_run_stats_t *stats, hipStream_t s) {
    const PlanHost *ph = nullptr;
    HC_TRY(get_plan(g, 0, g->n, tuning().pull, &ph));
    PrWs *ws = nullptr;
    HC_TRY(get_pr_ws(g, ph->dev.cnt[0], ph->dev.num_hubs, &ws));
    GraphView<OffT> v = view<OffT>(g);
    const int32_t ns = std::min<int32_t>(tuning().stage_pull, g->n);
    const int grid = resident_grid(reinterpret_cast<const void *>(&pr_step_kernel<OffT>), kBlock,
                                   ns * sizeof(float), grid_blocks());
    const int init_grid = std::max(1, std::min(grid, (g->n + kBlock - 1) / kBlock));
    const float base = (1.0f - damp) / (float)g->n;
    pr_init_kernel<OffT><<<init_grid, kBlock, 0, s>>>(
        v, ws->s[0], ws->c[0], ws->inv_deg, ws->hub_cnt, ph->dev.num_hubs, ws->ctl,
        max_iters, tol, (double)pr_epoch_bytes(g, ph), rule);
    HC_HIP(hipGetLastError());
    PrBufs b{ws->inv_deg, ws->hub_hi, ws->hub_lo, ws->hub_cnt, ws->hub_resid, ws->partials};
    hipEvent_t e0, e1;
    HC_HIP(hipEventCreate(&e0));
    HC_HIP(hipEventCreate(&e1));
    HC_HIP(hipEventRecord(e0, s));
    ws->kt.reset();
    for (int i = 0; i < max_iters; i++) {
        timed_launch(&ws->kt, pr_step_kernel<OffT>, dim3(grid), dim3(kBlock),
                     uint32_t(ns * sizeof(float)), s, ph->dev, v, ws->s[0], ws->s[1], ws->c[0],
                     ws->c[1], b, ws->ctl, base, damp, ns);
        pr_ctl_kernel<<<1, kBlock, 0, s>>>(ph->dev, ws->s[0], ws->s[1], ws->c[0],
                                           ws->c[1], b, ws->ctl, base, damp, grid);
    }
    HC_HIP(hipGetLastError());
    HC_HIP(hipEventRecord(e1, s));
    HC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    HC_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    PrCtl c;
    HC_HIP(hipMemcpyAsync(&c, ws->ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
    HC_HIP(hipStreamSynchronize(s));
    const float *final_s = (c.iter & 1) ? ws->s[1] : ws->s[0];
    HC_HIP(hipMemcpyAsync(scores, final_s, sizeof(float) * g->n,
                          hipMemcpyDeviceToDevice, s));
    HC_HIP(hipStreamSynchronize(s));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        (void)ms;
        stats->ms = double(c.t_end - c.t_start) / device_state().clock_khz;
        stats->epochs = c.iter;
        stats->pull_epochs = c.iter;
        stats->launches = max_iters;
        stats->reached = -1;
        stats->residual = c.residual;
        stats->alg_bytes = c.alg_bytes;
        if (ws->kt.on) {
            stats->kernel_launches = std::min(c.iter, ws->kt.pairs());
            stats->kernel_ms = ws->kt.sum_ms(stats->kernel_launches);
        }
    }
    return HC_OK;
}

// pr_pull_gpu's host loop: one mapped epoch per iteration, its stopping
// measure read back, until it says stop or max_iters ran.  BUCKETED runs
// the plan sweep of the epoch API.
template <typename OffT>
static int run_pr_pull(hc_graph_s *g, int mapping, const float *init, float *scores, float damp,
                       double tol, int max_iters, int rule, int blocks, int threads,
                       hc_run_stats_t *stats, hipStream_t s) {
    const PlanHost *ph = nullptr;
    if (mapping == HC_MAP_BUCKETED) HC_TRY(get_plan(g, 0, g->n, tuning().pull, &ph));
    PrWs *ws = nullptr;
    HC_TRY(get_pr_ws(g, ph ? ph->dev.cnt[0] : 0, ph ? ph->dev.num_hubs : 0, &ws));
    GraphView<OffT> v = view<OffT>(g);
    const int32_t ns = std::min<int32_t>(tuning().stage_pull, g->n);
    int nb = blocks, nt = threads;
    if (mapping == HC_MAP_BUCKETED) {
        nb = std::max(1, std::min(std::max(ph->total_units, 1), grid_blocks()));
        nt = kBlock;
    }
    double *dmem = nullptr;  // nb partials | the iteration's total
    HC_HIP(hipMalloc(&dmem, sizeof(double) * (nb + 1)));
    double *partials = dmem, *total = dmem + nb;
    const int init_grid = std::max(1, std::min(grid_blocks(), (g->n + kBlock - 1) / kBlock));
    pr_pull_init_kernel<OffT><<<init_grid, kBlock, 0, s>>>(v, init, ws->s[0], ws->c[0],
                                                           ws->inv_deg);
    int rc = hipGetLastError() == hipSuccess ? HC_OK : fail(HC_ERR_HIP, "PageRank init failed");
    const float base = (1.0f - damp) / (float)g->n;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == HC_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
                        hipEventRecord(e0, s) != hipSuccess))
        rc = fail(HC_ERR_HIP, "event setup failed");
    int it = 0;
    double measure = 0.0;
    while (rc == HC_OK && it < max_iters) {
        const int a = it & 1;
        PrEpoch ep{ws->s[a], ws->c[a], ws->s[a ^ 1], ws->c[a ^ 1], base, damp, rule};
        switch (mapping) {
            case HC_MAP_THREAD:
                pr_pull_mapped_kernel<HC_MAP_THREAD, OffT><<<nb, nt, 0, s>>>(v, ep, ws->inv_deg,
                                                                             partials);
                break;
            case HC_MAP_WAVE:
                pr_pull_mapped_kernel<HC_MAP_WAVE, OffT><<<nb, nt, 0, s>>>(v, ep, ws->inv_deg,
                                                                           partials);
                break;
            case HC_MAP_BLOCK:
                pr_pull_mapped_kernel<HC_MAP_BLOCK, OffT><<<nb, nt, 0, s>>>(v, ep, ws->inv_deg,
                                                                            partials);
                break;
            default: {
                PrBufs b{ws->inv_deg, ws->hub_hi, ws->hub_lo, ws->hub_cnt, ws->hub_resid, partials};
                pr_epoch_kernel<OffT><<<nb, kBlock, ns * sizeof(float), s>>>(ph->dev, v, ep, b, ns);
                break;
            }
        }
        if (mapping == HC_MAP_BUCKETED) {
            PrBufs b{ws->inv_deg, ws->hub_hi, ws->hub_lo, ws->hub_cnt, ws->hub_resid, partials};
            pr_pull_zero_kernel<<<1, 64, 0, s>>>(total);
            pr_epoch_finish_kernel<<<1, kBlock, 0, s>>>(ph->dev, ep, b, nb, total);
        } else {
            pr_pull_total_kernel<<<1, kBlock, 0, s>>>(partials, nb, total);
        }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&measure, total, sizeof(double), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(HC_ERR_HIP, "PageRank epoch %d failed", it);
            break;
        }
        it++;
        // rule 0: FP64 L1 residual below tol; rule 1: no vertex moved.
        if (rule ? measure == 0.0 : measure < tol) break;
    }
    float ms = 0.0f;
    if (rc == HC_OK && (hipEventRecord(e1, s) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(HC_ERR_HIP, "event timing failed");
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (rc == HC_OK &&
        (hipMemcpyAsync(scores, ws->s[it & 1], sizeof(float) * g->n, hipMemcpyDeviceToDevice,
                        s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        rc = fail(HC_ERR_HIP, "score copy failed");
    hipFree(dmem);
    if (rc == HC_OK && stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->ms = ms;
        stats->epochs = it;
        stats->pull_epochs = it;
        stats->launches = it;
        stats->reached = -1;
        stats->residual = measure;
    }
    return rc;
}

}  // namespace hc

using namespace hc;

extern "C" {

int hc_pr_pull(hc_graph_t g, int mapping, const float *init_score, float *scores, int blocks,
               int threads, const hc_pr_opts_t *opts, hc_run_stats_t *stats,
               hc_stream_t stream) {
    HC_CHECK_ARG(g && scores, "NULL argument");
    HC_CHECK_ARG(g->row0 == 0 && g->rows == g->n, "full runs need a whole-graph handle");
    HC_CHECK_ARG(g->n > 0, "empty graph");
    HC_CHECK_ARG(mapping >= HC_MAP_THREAD && mapping <= HC_MAP_BUCKETED, "unknown mapping %d",
                 mapping);
    HC_CHECK_ARG(blocks >= 0 && threads >= 0 && threads <= 1024,
                 "launch shape must be >= 0 blocks and 0..1024 threads");
    if (blocks == 0) blocks = 64;
    if (threads == 0) threads = 1024;
    if (mapping == HC_MAP_WAVE || mapping == HC_MAP_BLOCK)
        HC_CHECK_ARG(threads % kWave == 0,
                     "thread count must be a multiple of 64 for this mapping");
    const float damp = opts && opts->damping > 0 ? opts->damping : 0.85f;
    const double tol = opts && opts->tolerance != 0 ? opts->tolerance : 1e-4;
    const int max_iters = opts && opts->max_iters > 0 ? opts->max_iters : 20;
    const int rule = opts ? opts->rule : 0;
    HC_CHECK_ARG(damp < 1.0f, "damping must be in (0,1)");
    HC_CHECK_ARG(rule == 0 || rule == 1, "rule must be 0 (L1 residual) or 1 (per vertex)");
    HC_CHECK_ARG(max_iters <= 100000, "max_iters too large");
    HC_TRY(ensure_init());
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    return g->off64 ? run_pr_pull<uint64_t>(g, mapping, init_score, scores, damp, tol,
                                            max_iters, rule, blocks, threads, stats, s)
                    : run_pr_pull<uint32_t>(g, mapping, init_score, scores, damp, tol,
                                            max_iters, rule, blocks, threads, stats, s);
}

int hc_pagerank(hc_graph_t g, float *scores, const hc_pr_opts_t *opts,
                hc_run_stats_t *stats, hc_stream_t stream) {
    HC_CHECK_ARG(g && scores, "NULL argument");
    HC_CHECK_ARG(g->row0 == 0 && g->rows == g->n,
                 "full runs need a whole-graph handle");
    HC_CHECK_ARG(g->n > 0, "empty graph");
    float damp = opts && opts->damping > 0 ? opts->damping : 0.85f;
    // tolerance 0 selects the default; a negative tolerance disables the
    // residual test (exactly max_iters iterations).
    double tol = opts && opts->tolerance != 0 ? opts->tolerance : 1e-4;
    int max_iters = opts && opts->max_iters > 0 ? opts->max_iters : 20;
    const int rule = opts ? opts->rule : 0;
    HC_CHECK_ARG(damp < 1.0f, "damping must be in (0,1)");
    HC_CHECK_ARG(rule == 0 || rule == 1, "rule must be 0 (L1 residual) or 1 (per vertex)");
    HC_CHECK_ARG(max_iters <= 100000, "max_iters too large");
    HC_TRY(ensure_init());
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    return g->off64 ? run_pr<uint64_t>(g, scores, damp, tol, max_iters, rule, stats, s)
                    : run_pr<uint32_t>(g, scores, damp, tol, max_iters, rule, stats, s);
}

int hc_pr_pull_epoch(hc_graph_t g, int32_t start_id, int32_t end_id,
                     float damping, const float *score_in,
                     const float *contrib, const double *inv_degree,
                     float *score_out, float *contrib_out,
                     double *residual_out, hc_stream_t stream) {
    HC_CHECK_ARG(g && score_in && contrib && inv_degree && score_out && contrib_out,
                 "NULL argument");
    HC_CHECK_ARG(g->row0 <= start_id && start_id <= end_id &&
                     end_id <= g->row0 + g->rows,
                 "range [%d,%d) outside the rows held [%d,%d)", start_id,
                 end_id, g->row0, g->row0 + g->rows);
    HC_CHECK_ARG(damping > 0 && damping < 1, "damping must be in (0,1)");
    HC_TRY(ensure_init());
    if (start_id == end_id) return HC_OK;
    GraphCall call(g);
    hipStream_t s = pick_stream(stream);
    const PlanHost *ph = nullptr;
    HC_TRY(get_plan(g, start_id, end_id, tuning().pull, &ph));
    PrWs *ws = nullptr;
    HC_TRY(get_pr_ws(g, ph->dev.cnt[0], ph->dev.num_hubs, &ws));
    PrBufs b{inv_degree, ws->hub_hi, ws->hub_lo, ws->hub_cnt, ws->hub_resid, ws->partials};
    PrEpoch ep{score_in, contrib, score_out, contrib_out,
               (1.0f - damping) / (float)g->n, damping, 0};
    const int nb = std::max(1, std::min(std::max(ph->total_units, 1), grid_blocks()));
    const int32_t ns = std::min<int32_t>(tuning().stage_pull, g->n);
    if (g->off64)
        pr_epoch_kernel<uint64_t><<<nb, kBlock, ns * sizeof(float), s>>>(
            ph->dev, view<uint64_t>(g), ep, b, ns);
    else
        pr_epoch_kernel<uint32_t><<<nb, kBlock, ns * sizeof(float), s>>>(
            ph->dev, view<uint32_t>(g), ep, b, ns);
    pr_epoch_finish_kernel<<<1, kBlock, 0, s>>>(ph->dev, ep, b, nb, residual_out);
    HC_HIP(hipGetLastError());
    return HC_OK;
}

}  // extern "C"
