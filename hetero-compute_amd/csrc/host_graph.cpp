// Host-side graph construction and I/O (compiled with g++ -fopenmp).
//
//  * Generator: R-MAT / uniform edge lists in 2^18-edge blocks, each block's
//    mt19937 reseeded with seed + block index, as the modified gapbs
//    generator does (src/gapbs/generator.h:53-115); int32 weights in [1,255]
//    from the same block reseeding (the int variant kept commented out at
//    src/gapbs/generator.h:133-147).
//  * Builder: symmetrize, drop self-loops and duplicates, relabel by
//    descending degree (src/gapbs/builder.h:140-172, 297-313, 366-406).  The
//    implementation is a two-level counting partition (vertex blocks, then
//    rows) so no step depends on thread timing: the output is a pure function
//    of the edge list.
//  * File I/O in the reference's binary layout (src/graph.cuh:207-239,
//    src/gapbs/gapbs.cuh:58-112), with validation the reference lacks.
//  * SourcePicker (src/graph.cuh:106-141) and compute_equal_edge_ranges
//    (src/graph.cuh:156-201).
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "hc_internal.h"

namespace hc {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

namespace {

constexpr int64_t kBlockSize = 1 << 18;  // generator.h:171

template <typename T>
T *alloc_array(size_t n) {
    if (n == 0) n = 1;
    return static_cast<T *>(std::malloc(n * sizeof(T)));
}

// Exclusive prefix sum of counts[0..n) into out[0..n], in parallel blocks.
void prefix_sum(const int64_t *counts, int64_t n, int64_t *out) {
    int nt = omp_get_max_threads();
    std::vector<int64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        int T = omp_get_num_threads();
        int64_t lo = n * t / T, hi = n * (t + 1) / T, s = 0;
        for (int64_t i = lo; i < hi; i++) s += counts[i];
        part[t + 1] = s;
#pragma omp barrier
#pragma omp single
        for (int i = 0; i < T; i++) part[i + 1] += part[i];
        int64_t run = part[t];
        for (int64_t i = lo; i < hi; i++) {
            out[i] = run;
            run += counts[i];
        }
    }
    out[n] = n > 0 ? out[n - 1] + counts[n - 1] : 0;
}

inline uint64_t pack_vw(int32_t v, int32_t w) {
    return (static_cast<uint64_t>(static_cast<uint32_t>(v)) << 32) |
           static_cast<uint32_t>(w);
}
inline int32_t unpack_v(uint64_t x) { return static_cast<int32_t>(x >> 32); }
inline int32_t unpack_w(uint64_t x) {
    return static_cast<int32_t>(x & 0xffffffffu);
}

// Sorts one row by (neighbour, weight), keeps the first (lightest) copy of
// each neighbour, drops `self` (builder.h:148-157 sort/unique/remove).
int64_t squish_row(uint64_t *row, int64_t len, int32_t self, bool drop_self) {
    std::sort(row, row + len);
    int64_t out = 0;
    for (int64_t i = 0; i < len; i++) {
        int32_t v = unpack_v(row[i]);
        if (drop_self && v == self) continue;
        if (out > 0 && unpack_v(row[out - 1]) == v) continue;
        row[out++] = row[i];
    }
    return out;
}

struct Csr {
    int32_t n = 0;
    std::vector<int64_t> deg;  // per vertex
    uint64_t *rows = nullptr;  // packed (v,w), rows back to back (prefix of deg)
    std::vector<int64_t> off;  // n+1
};

}  // namespace

// Builds a squished (optionally symmetrized and relabelled) CSR from an edge
// list.  Everything is order-independent: rows are sorted by (v, w) before
// duplicates are removed, so the result does not depend on scatter order.
int build_csr(int32_t n, int64_t ne, const int32_t *src, const int32_t *dst,
              const int32_t *wts, bool symmetrize, bool relabel,
              bool weighted, hc_host_graph_t *out) {
    if (n < 0) return fail(HC_ERR_INVALID, "num_nodes must be >= 0");
    if (ne < 0) return fail(HC_ERR_INVALID, "num_edges must be >= 0");
    // Validate endpoints.
    int64_t bad = -1;
#pragma omp parallel for reduction(max : bad)
    for (int64_t e = 0; e < ne; e++) {
        if (src[e] < 0 || src[e] >= n || dst[e] < 0 || dst[e] >= n)
            bad = std::max(bad, e);
    }
    if (bad >= 0)
        return fail(HC_ERR_INVALID, "edge %lld has an endpoint outside [0,%d)",
                    (long long)bad, n);
    if (wts) {
#pragma omp parallel for reduction(max : bad)
        for (int64_t e = 0; e < ne; e++)
            if (wts[e] < 0 || wts[e] >= HC_INF_DIST) bad = std::max(bad, e);
        if (bad >= 0)
            return fail(HC_ERR_INVALID,
                        "edge %lld has a weight outside [0, 2^30)",
                        (long long)bad);
    }

    const int64_t D = symmetrize ? 2 * ne : ne;
    // ---- Level 1: partition directed edges by vertex block of the source.
    int pshift = 0;
    while (((int64_t)(n > 0 ? n - 1 : 0) >> pshift) >= 4096) pshift++;
    const int64_t P = n > 0 ? ((int64_t)(n - 1) >> pshift) + 1 : 1;
    const int T = omp_get_max_threads();
    std::vector<int64_t> hist((size_t)T * P, 0);
#pragma omp parallel num_threads(T)
    {
        int t = omp_get_thread_num();
        int TT = omp_get_num_threads();
        for (int tt = t; tt < T; tt += TT) {
            int64_t lo = ne * tt / T, hi = ne * (tt + 1) / T;
            int64_t *hh = &hist[(size_t)tt * P];
            for (int64_t e = lo; e < hi; e++) {
                hh[src[e] >> pshift]++;
                if (symmetrize) hh[dst[e] >> pshift]++;
            }
        }
    }
    // Positions: partition-major, slice-minor.
    std::vector<int64_t> pstart(P + 1, 0);
    {
        int64_t run = 0;
        for (int64_t p = 0; p < P; p++) {
            pstart[p] = run;
            for (int t = 0; t < T; t++) {
                int64_t c = hist[(size_t)t * P + p];
                hist[(size_t)t * P + p] = run;
                run += c;
            }
        }
        pstart[P] = run;
    }
    int32_t *eu = alloc_array<int32_t>(D);
    uint64_t *evw = alloc_array<uint64_t>(D);
    if (!eu || !evw) {
        std::free(eu);
        std::free(evw);
        return fail(HC_ERR_OOM, "host allocation of %lld edges failed",
                    (long long)D);
    }
#pragma omp parallel num_threads(T)
    {
        int t = omp_get_thread_num();
        int TT = omp_get_num_threads();
        for (int tt = t; tt < T; tt += TT) {
            int64_t lo = ne * tt / T, hi = ne * (tt + 1) / T;
            int64_t *cur = &hist[(size_t)tt * P];
            for (int64_t e = lo; e < hi; e++) {
                int32_t u = src[e], v = dst[e];
                int32_t w = wts ? wts[e] : 0;
                int64_t pos = cur[u >> pshift]++;
                eu[pos] = u;
                evw[pos] = pack_vw(v, w);
                if (symmetrize) {
                    pos = cur[v >> pshift]++;
                    eu[pos] = v;
                    evw[pos] = pack_vw(u, w);
                }
            }
        }
    }
    std::vector<int64_t>().swap(hist);

    // ---- Level 2: within each vertex block, group by row, squish rows.
    std::vector<int64_t> deg(n > 0 ? n : 1, 0);
    uint64_t *rowbuf = alloc_array<uint64_t>(D);
    if (!rowbuf) {
        std::free(eu);
        std::free(evw);
        return fail(HC_ERR_OOM, "host allocation failed");
    }
#pragma omp parallel
    {
        std::vector<int64_t> cnt, cur;
#pragma omp for schedule(dynamic, 1)
        for (int64_t p = 0; p < P; p++) {
            int64_t a = pstart[p], b = pstart[p + 1];
            int32_t v0 = (int32_t)(p << pshift);
            int32_t v1 = (int32_t)std::min<int64_t>((p + 1) << pshift, n);
            int32_t nv = v1 - v0;
            cnt.assign(nv + 1, 0);
            for (int64_t i = a; i < b; i++) cnt[eu[i] - v0]++;
            cur.assign(nv + 1, 0);
            int64_t run = a;
            for (int32_t i = 0; i < nv; i++) {
                cur[i] = run;
                run += cnt[i];
            }
            for (int64_t i = a; i < b; i++) rowbuf[cur[eu[i] - v0]++] = evw[i];
            // Rows now at rowbuf[start_i, start_i+cnt_i); squish each and
            // compact them to the front of the block's region.
            int64_t rs = a, wpos = a;
            for (int32_t i = 0; i < nv; i++) {
                int64_t len = squish_row(rowbuf + rs, cnt[i], v0 + i, true);
                if (wpos != rs)
                    std::memmove(rowbuf + wpos, rowbuf + rs,
                                 len * sizeof(uint64_t));
                deg[v0 + i] = len;
                wpos += len;
                rs += cnt[i];
            }
            // Record block's compacted length in pstart-space via evw reuse.
        }
    }
    std::free(eu);
    std::free(evw);

    // Squished rows live at rowbuf[pstart[p] + local prefix]; gather them
    // into a dense CSR.
    std::vector<int64_t> off(n + 1, 0);
    prefix_sum(deg.data(), n, off.data());
    const int64_t M = off[n];
    uint64_t *rows = alloc_array<uint64_t>(M);
    if (!rows) {
        std::free(rowbuf);
        return fail(HC_ERR_OOM, "host allocation failed");
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < P; p++) {
        int32_t v0 = (int32_t)(p << pshift);
        int32_t v1 = (int32_t)std::min<int64_t>((p + 1) << pshift, n);
        int64_t src_pos = pstart[p];
        for (int32_t v = v0; v < v1; v++) {
            std::memcpy(rows + off[v], rowbuf + src_pos,
                        deg[v] * sizeof(uint64_t));
            src_pos += deg[v];
        }
    }
    std::free(rowbuf);

    // ---- Relabel by descending degree (builder.h:366-406).
    if (relabel && n > 0) {
        std::vector<std::pair<int64_t, int32_t>> dp(n);
#pragma omp parallel for
        for (int32_t v = 0; v < n; v++) dp[v] = std::make_pair(deg[v], v);
        std::sort(dp.begin(), dp.end(),
                  std::greater<std::pair<int64_t, int32_t>>());
        std::vector<int32_t> new_ids(n);
        std::vector<int64_t> ndeg(n);
#pragma omp parallel for
        for (int32_t i = 0; i < n; i++) {
            ndeg[i] = dp[i].first;
            new_ids[dp[i].second] = i;
        }
        std::vector<int64_t> noff(n + 1, 0);
        prefix_sum(ndeg.data(), n, noff.data());
        uint64_t *nrows = alloc_array<uint64_t>(M);
        if (!nrows) {
            std::free(rows);
            return fail(HC_ERR_OOM, "host allocation failed");
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int32_t u = 0; u < n; u++) {
            int64_t d = deg[u];
            uint64_t *dstrow = nrows + noff[new_ids[u]];
            const uint64_t *srow = rows + off[u];
            for (int64_t i = 0; i < d; i++)
                dstrow[i] = pack_vw(new_ids[unpack_v(srow[i])],
                                    unpack_w(srow[i]));
            std::sort(dstrow, dstrow + d);
        }
        std::free(rows);
        rows = nrows;
        off.swap(noff);
    }

    // ---- Emit structure-of-arrays CSR.
    out->num_nodes = n;
    out->num_edges = M;
    out->index = alloc_array<int64_t>(n + 1);
    out->neighbors = alloc_array<int32_t>(M);
    out->weights = weighted ? alloc_array<int32_t>(M) : nullptr;
    if (!out->index || !out->neighbors || (weighted && !out->weights)) {
        std::free(out->index);
        std::free(out->neighbors);
        std::free(out->weights);
        std::free(rows);
        out->index = nullptr;
        out->neighbors = out->weights = nullptr;
        return fail(HC_ERR_OOM, "host allocation failed");
    }
    std::memcpy(out->index, off.data(), (n + 1) * sizeof(int64_t));
#pragma omp parallel for
    for (int64_t i = 0; i < M; i++) {
        out->neighbors[i] = unpack_v(rows[i]);
        if (weighted) out->weights[i] = unpack_w(rows[i]);
    }
    std::free(rows);
    return HC_OK;
}

int generate_edges(bool uniform, int scale, int degree, bool weighted,
                   int64_t seed, std::vector<int32_t> &src,
                   std::vector<int32_t> &dst, std::vector<int32_t> &wts) {
    if (scale < 1 || scale > 30)
        return fail(HC_ERR_INVALID, "scale must be in [1,30], got %d", scale);
    if (degree < 1 || degree > 1024)
        return fail(HC_ERR_INVALID, "degree must be in [1,1024], got %d",
                    degree);
    const int64_t n = int64_t(1) << scale;
    const int64_t ne = n * degree;
    try {
        src.resize(ne);
        dst.resize(ne);
        if (weighted) wts.resize(ne);
    } catch (const std::bad_alloc &) {
        return fail(HC_ERR_OOM, "edge list of %lld edges does not fit",
                    (long long)ne);
    }
    const int64_t nblocks = (ne + kBlockSize - 1) / kBlockSize;
    if (uniform) {
        // MakeUniformEL (generator.h:65-80); u is drawn before v.
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t blk = 0; blk < nblocks; blk++) {
            std::mt19937 rng;
            std::uniform_int_distribution<int32_t> udist(0, (int32_t)(n - 1));
            rng.seed((uint32_t)(seed + blk));
            int64_t hi = std::min(ne, (blk + 1) * kBlockSize);
            for (int64_t e = blk * kBlockSize; e < hi; e++) {
                int32_t u = udist(rng);
                int32_t v = udist(rng);
                src[e] = u;
                dst[e] = v;
            }
        }
    } else {
        // MakeRMatEL (generator.h:82-115) then PermuteIDs (:53-63).
        const float A = 0.57f, B = 0.19f, C = 0.19f;
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t blk = 0; blk < nblocks; blk++) {
            std::mt19937 rng;
            std::uniform_real_distribution<float> udist(0, 1.0f);
            rng.seed((uint32_t)(seed + blk));
            int64_t hi = std::min(ne, (blk + 1) * kBlockSize);
            for (int64_t e = blk * kBlockSize; e < hi; e++) {
                int32_t s = 0, d = 0;
                for (int depth = 0; depth < scale; depth++) {
                    float r = udist(rng);
                    s <<= 1;
                    d <<= 1;
                    if (r < A + B) {
                        if (r > A) d++;
                    } else {
                        s++;
                        if (r > A + B + C) d++;
                    }
                }
                src[e] = s;
                dst[e] = d;
            }
        }
        std::vector<int32_t> perm(n);
        for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
        std::mt19937 prng((uint32_t)seed);
        std::shuffle(perm.begin(), perm.end(), prng);
#pragma omp parallel for
        for (int64_t e = 0; e < ne; e++) {
            src[e] = perm[src[e]];
            dst[e] = perm[dst[e]];
        }
    }
    if (weighted) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t blk = 0; blk < nblocks; blk++) {
            std::mt19937 rng;
            std::uniform_int_distribution<int> udist(1, 255);
            rng.seed((uint32_t)(seed + blk));
            int64_t hi = std::min(ne, (blk + 1) * kBlockSize);
            for (int64_t e = blk * kBlockSize; e < hi; e++)
                wts[e] = udist(rng);
        }
    }
    return HC_OK;
}

}  // namespace hc

using namespace hc;

extern "C" {

const char *hc_last_error(void) { return g_last_error.c_str(); }

int hc_version(void) { return 10000; }

int hc_host_graph_generate(int uniform, int scale, int degree, int weighted,
                           int64_t seed, hc_host_graph_t *out) {
    HC_CHECK_ARG(out, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    std::vector<int32_t> src, dst, wts;
    HC_TRY(generate_edges(uniform != 0, scale, degree, weighted != 0, seed,
                          src, dst, wts));
    const int32_t n = (int32_t)(int64_t(1) << scale);
    return build_csr(n, (int64_t)src.size(), src.data(), dst.data(),
                     weighted ? wts.data() : nullptr, true, true,
                     weighted != 0, out);
}

int hc_host_graph_from_edges(const int32_t *src, const int32_t *dst,
                             const int32_t *weights, int64_t num_edges,
                             int32_t num_nodes, int symmetrize, int relabel,
                             hc_host_graph_t *out) {
    HC_CHECK_ARG(out, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    HC_CHECK_ARG(num_edges == 0 || (src && dst), "edge arrays are NULL");
    return build_csr(num_nodes, num_edges, src, dst, weights, symmetrize != 0,
                     relabel != 0, weights != nullptr, out);
}

void hc_host_graph_free(hc_host_graph_t *g) {
    if (!g) return;
    std::free(g->index);
    std::free(g->neighbors);
    std::free(g->weights);
    g->index = nullptr;
    g->neighbors = nullptr;
    g->weights = nullptr;
    g->num_nodes = 0;
    g->num_edges = 0;
}

// operator>> (src/graph.cuh:207-224) plus the validation it lacks.
int hc_host_graph_read(const char *path, int weighted, hc_host_graph_t *out) {
    HC_CHECK_ARG(path && out, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    std::ifstream ifs(path, std::ios::in | std::ios::binary);
    if (!ifs) return fail(HC_ERR_IO, "cannot open %s: %s", path, strerror(errno));
    int32_t n = 0;
    int64_t m = 0;
    ifs.read(reinterpret_cast<char *>(&n), sizeof(n));
    ifs.read(reinterpret_cast<char *>(&m), sizeof(m));
    if (!ifs) return fail(HC_ERR_IO, "%s: truncated header", path);
    if (n < 0 || m < 0)
        return fail(HC_ERR_INVALID, "%s: bad header (N=%d, m=%lld)", path, n,
                    (long long)m);
    ifs.seekg(0, std::ios::end);
    const int64_t fsize = (int64_t)ifs.tellg();
    const int64_t esize = weighted ? 8 : 4;
    const int64_t need = 12 + (int64_t(n) + 1) * 8 + m * esize;
    if (fsize < need)
        return fail(HC_ERR_INVALID, "%s: file has %lld bytes, header needs %lld",
                    path, (long long)fsize, (long long)need);
    ifs.seekg(12, std::ios::beg);
    hc_host_graph_t g{};
    g.num_nodes = n;
    g.num_edges = m;
    g.index = alloc_array<int64_t>(int64_t(n) + 1);
    g.neighbors = alloc_array<int32_t>(m);
    g.weights = weighted ? alloc_array<int32_t>(m) : nullptr;
    if (!g.index || !g.neighbors || (weighted && !g.weights)) {
        hc_host_graph_free(&g);
        return fail(HC_ERR_OOM, "host allocation failed");
    }
    ifs.read(reinterpret_cast<char *>(g.index), (int64_t(n) + 1) * 8);
    if (g.index[0] != 0 || g.index[n] != m) {
        hc_host_graph_free(&g);
        return fail(HC_ERR_INVALID, "%s: index[0]/index[N] do not match m",
                    path);
    }
    for (int32_t i = 0; i < n; i++)
        if (g.index[i + 1] < g.index[i]) {
            hc_host_graph_free(&g);
            return fail(HC_ERR_INVALID, "%s: index not monotone at %d", path,
                        i);
        }
    if (weighted) {
        const int64_t chunk = 1 << 20;
        std::vector<char> buf(chunk * 8);
        for (int64_t e = 0; e < m; e += chunk) {
            int64_t c = std::min(chunk, m - e);
            ifs.read(buf.data(), c * 8);
            for (int64_t i = 0; i < c; i++) {
                int32_t v;
                float w;
                std::memcpy(&v, buf.data() + 8 * i, 4);
                std::memcpy(&w, buf.data() + 8 * i + 4, 4);
                if (!(w >= 0.0f) || w >= 1073741824.0f || std::floor(w) != w) {
                    hc_host_graph_free(&g);
                    return fail(HC_ERR_UNSUPPORTED,
                                "%s: edge %lld weight %g is not an integer in "
                                "[0,2^30) (int32 distances)",
                                path, (long long)(e + i), (double)w);
                }
                g.neighbors[e + i] = v;
                g.weights[e + i] = (int32_t)w;
            }
        }
    } else {
        ifs.read(reinterpret_cast<char *>(g.neighbors), m * 4);
    }
    if (!ifs) {
        hc_host_graph_free(&g);
        return fail(HC_ERR_IO, "%s: short read", path);
    }
    for (int64_t e = 0; e < m; e++)
        if (g.neighbors[e] < 0 || g.neighbors[e] >= n) {
            hc_host_graph_free(&g);
            return fail(HC_ERR_INVALID, "%s: neighbour %lld out of range", path,
                        (long long)e);
        }
    *out = g;
    return HC_OK;
}

// Serializers (src/gapbs/gapbs.cuh:66-112).  The unweighted layout writes
// exactly m int32 ids (the reference writes m*8 bytes from a 4-byte array,
// gapbs.cuh:79-80; readers of either ignore trailing bytes).
int hc_host_graph_write(const char *path, const hc_host_graph_t *g,
                        int weighted) {
    HC_CHECK_ARG(path && g && g->index, "NULL argument");
    HC_CHECK_ARG(!weighted || g->weights, "graph has no weights");
    std::ofstream ofs(path, std::ios::out | std::ios::binary);
    if (!ofs) return fail(HC_ERR_IO, "cannot create %s", path);
    ofs.write(reinterpret_cast<const char *>(&g->num_nodes), 4);
    ofs.write(reinterpret_cast<const char *>(&g->num_edges), 8);
    ofs.write(reinterpret_cast<const char *>(g->index),
              (int64_t(g->num_nodes) + 1) * 8);
    if (weighted) {
        const int64_t chunk = 1 << 20;
        std::vector<char> buf(chunk * 8);
        for (int64_t e = 0; e < g->num_edges; e += chunk) {
            int64_t c = std::min(chunk, g->num_edges - e);
            for (int64_t i = 0; i < c; i++) {
                float w = (float)g->weights[e + i];
                std::memcpy(buf.data() + 8 * i, &g->neighbors[e + i], 4);
                std::memcpy(buf.data() + 8 * i + 4, &w, 4);
            }
            ofs.write(buf.data(), c * 8);
        }
    } else {
        ofs.write(reinterpret_cast<const char *>(g->neighbors),
                  g->num_edges * 4);
    }
    if (!ofs) return fail(HC_ERR_IO, "write to %s failed", path);
    return HC_OK;
}

int hc_pick_sources(const hc_host_graph_t *g, int64_t seed, int32_t count,
                    int32_t *out) {
    HC_CHECK_ARG(g && g->index && out, "NULL argument");
    HC_CHECK_ARG(count >= 0, "count must be >= 0");
    const int32_t n = g->num_nodes;
    auto deg = [&](int32_t v) { return g->index[v + 1] - g->index[v]; };
    auto self_cycle = [&](int32_t v) {
        for (int64_t i = g->index[v]; i < g->index[v + 1]; i++)
            if (g->neighbors[i] == v) return true;
        return false;
    };
    bool any = false;
    for (int32_t v = 0; v < n && !any; v++)
        if (deg(v) > 1 || (deg(v) == 1 && !self_cycle(v))) any = true;
    if (!any)
        return fail(HC_ERR_INVALID,
                    "graph has no vertex a source can be picked from");
    std::mt19937 rng((uint32_t)seed);
    std::uniform_int_distribution<int32_t> udist(0, n - 1);
    for (int32_t i = 0; i < count; i++) {
        int32_t s;
        do {
            s = udist(rng);
        } while (deg(s) == 0 || (deg(s) == 1 && self_cycle(s)));
        out[i] = s;
    }
    return HC_OK;
}

// compute_equal_edge_ranges (src/graph.cuh:156-201).  Two defects of the
// reference are closed: boundaries are never written past out[S], and the
// last boundary is always N, so segments always cover [0, N).
int hc_equal_edge_ranges(const hc_host_graph_t *g, int32_t num_segments,
                         int32_t chunk_size, int32_t *out) {
    HC_CHECK_ARG(g && g->index && out, "NULL argument");
    HC_CHECK_ARG(num_segments >= 1, "num_segments must be >= 1");
    HC_CHECK_ARG(chunk_size >= 1, "chunk_size must be >= 1");
    const int32_t n = g->num_nodes;
    const int64_t avg = g->num_edges / num_segments;
    for (int32_t i = 0; i <= num_segments; i++) out[i] = n;
    out[0] = 0;
    int32_t end_id = 0, seg_id = 0;
    int64_t seg_deg = 0;
    while (end_id != n) {
        seg_deg += g->index[end_id + 1] - g->index[end_id];
        if (seg_deg >= avg && seg_id + 1 < num_segments) {
            out[seg_id + 1] = end_id;
            seg_deg = 0;
            seg_id++;
        }
        end_id++;
    }
    out[num_segments] = n;
    for (int32_t i = 1; i < num_segments; i++) {
        int32_t diff = out[i] % chunk_size;
        if (diff != 0) {
            if (diff < chunk_size / 2)
                out[i] = out[i] / chunk_size * chunk_size;
            else
                out[i] = std::min(n, (out[i] / chunk_size + 1) * chunk_size);
        }
    }
    for (int32_t i = 1; i <= num_segments; i++)
        out[i] = std::max(out[i], out[i - 1]);
    return HC_OK;
}

}  // extern "C"
