// Runtime: device selection, memory and stream helpers, CSR upload into HBM
// (replacing copy_subgraph_to_device, src/graph.cuh:250-286) and the
// degree-bucket plans the fused epoch kernels run from.
#include <algorithm>
#include <cstring>

#include "device_common.h"
#include "frontier.h"

namespace hc {

DeviceState &device_state() {
    static DeviceState s;
    return s;
}

int ensure_init() {
    DeviceState &s = device_state();
    if (s.ready) {
        HC_HIP(hipSetDevice(s.device));
        return HC_OK;
    }
    return hc_init(0);
}

Tuning &global_tuning() {
    static Tuning t;
    return t;
}

namespace {
thread_local Tuning *tls_tuning = nullptr;
}  // namespace

Tuning &tuning() { return tls_tuning ? *tls_tuning : global_tuning(); }

TuningScope::TuningScope(Tuning *t) : prev_(tls_tuning), set_(t != nullptr) {
    if (set_) tls_tuning = t;
}

TuningScope::~TuningScope() {
    if (set_) tls_tuning = prev_;
}

Tuning *GraphCall::effective(hc_graph_s *g, Tuning *eff) {
    GraphTuning &gt = *g->tune;
    std::lock_guard<std::mutex> lk(gt.mu);
    if (gt.over.empty()) return nullptr;
    *eff = global_tuning();
    for (const auto &kv : gt.over) (void)set_knob(*eff, kv.first.c_str(), kv.second);
    return eff;
}

GraphCall::GraphCall(hc_graph_s *g) : lk_(g->mu), eff_(), scope_(effective(g, &eff_)) {}

int grid_blocks(int per_cu) {
    if (per_cu <= 0) per_cu = tuning().grid_per_cu;
    return std::max(1, device_state().cus * per_cu);
}

int resident_grid(const void *kernel, int block, size_t lds_bytes, int cap) {
    static std::mutex mu;
    static std::map<std::tuple<const void *, int, size_t>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(kernel, block, lds_bytes);
    auto it = cache.find(key);
    int per_cu = 0;
    if (it != cache.end()) {
        per_cu = it->second;
    } else {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds_bytes) !=
                hipSuccess ||
            per_cu <= 0)
            per_cu = 1;
        cache.emplace(key, per_cu);
    }
    return std::max(1, std::min(cap, device_state().cus * per_cu));
}

namespace {
int64_t g_knob_tmp;
int64_t *knob_ref(Tuning &t, const char *key, bool *is_int, int **iref) {
    *is_int = false;
    *iref = nullptr;
    struct {
        const char *k;
        int64_t *p;
    } i64[] = {
        {"sssp.block_threshold", &t.sssp_block_threshold},
        {"sssp.wave_threshold", &t.sssp_wave_threshold},
        {"sssp.g16_threshold", &t.sssp_g16_threshold},
        {"sssp.g4_threshold", &t.sssp_g4_threshold},
        {"pull.hub_threshold", &t.pull.hub_threshold},
        {"pull.hub_chunk", &t.pull.hub_chunk},
        {"pull.block_threshold", &t.pull.block_threshold},
        {"pull.wave_threshold", &t.pull.wave_threshold},
        {"pull.g16_threshold", &t.pull.g16_threshold},
        {"pull.g4_threshold", &t.pull.g4_threshold},
        {"bu.wave_threshold", &t.bu.wave_threshold},
        {"bu.g16_threshold", &t.bu.g16_threshold},
        {"bu.g4_threshold", &t.bu.g4_threshold},
        {"multi.hub_threshold", &t.multi.hub_threshold},
        {"multi.hub_chunk", &t.multi.hub_chunk},
        {"multi.wave_threshold", &t.multi.wave_threshold},
        {"multi.g16_threshold", &t.multi.g16_threshold},
        {"cb.min_slice", &t.cb_min_slice},
        {"cb.g4_slice", &t.cb_g4_slice},
        {"cb.g16_slice", &t.cb_g16_slice},
        {"cb.wave_slice", &t.cb_wave_slice},
        {"cb.block_slice", &t.cb_block_slice},
    };
    for (auto &x : i64)
        if (!strcmp(key, x.k)) return x.p;
    struct {
        const char *k;
        int *p;
    } i32[] = {{"grid_per_cu", &t.grid_per_cu},
               {"sssp.alpha", &t.sssp_alpha},
               {"sssp.filter", &t.sssp_filter},
               {"sssp.dist16", &t.sssp_dist16},
               {"sssp.push_dedup", &t.sssp_push_dedup},
               {"sssp.stage", &t.sssp_stage},
               {"sssp.grid_per_cu", &t.sssp_grid_per_cu},
               {"pull.stage", &t.stage_pull},
               {"bu.stage", &t.stage_bu},
               {"bu.remaining", &t.bu_remaining},
               {"push.chunk", &t.push_chunk},
               {"push.stage", &t.push_stage},
               {"bu.first", &t.bu_first},
               {"compact.grid", &t.compact_grid},
               {"bu.stage_min", &t.bu_stage_min},
               {"loop.batch", &t.loop_batch},
               {"profile.kernel_events", &t.kernel_events},
               {"multi.graph", &t.multi_graph},
               {"multi.grid_per_cu", &t.multi_grid_per_cu},
               {"multi.block", &t.multi_block},
               {"multi.delta", &t.multi_delta},
               {"sssp.delta", &t.sssp_delta},
               {"multi.push_div", &t.multi_push_div},
               {"msbfs.on", &t.msbfs},
               {"msbfs.alpha", &t.msbfs_alpha},
               {"msbfs.col_shift", &t.msbfs_col_shift},
               {"many.block", &t.many_block},
               {"many.per_cu", &t.many_per_cu},
               {"many.light_deg", &t.many_light_deg},
               {"many.heavy_cap", &t.many_heavy_cap},
               {"many.lds_nodes", &t.many_lds_nodes},
               {"pull.col_shift", &t.pull_col_shift}};
    for (auto &x : i32)
        if (!strcmp(key, x.k)) {
            *is_int = true;
            *iref = x.p;
            return &g_knob_tmp;
        }
    return nullptr;
}
}  // namespace

hipStream_t pick_stream(hc_stream_t s) {
    return s ? static_cast<hipStream_t>(s) : device_state().stream;
}

// Builds (or returns the cached) plan for rows [start, end).
int get_plan(hc_graph_s *g, int32_t start, int32_t end, const PlanConfig &cfg,
             const PlanHost **out) {
    auto key = std::make_tuple(start, end, cfg);
    auto it = g->plans.find(key);
    if (it != g->plans.end()) {
        *out = &it->second;
        return HC_OK;
    }
    auto deg = [&](int32_t u) {
        return g->h_off[u - g->row0 + 1] - g->h_off[u - g->row0];
    };
    const int64_t dmin = std::max<int64_t>(cfg.min_degree, 0), dmax = cfg.max_degree;
    auto cls_of = [&](int64_t d) {
        if (d == 0) return -1;
        if (d > cfg.hub_threshold) return 0;
        if (d > cfg.block_threshold) return 1;
        if (d > cfg.wave_threshold) return 2;
        if (d > cfg.g16_threshold) return 3;
        if (d > cfg.g4_threshold) return 4;
        return 5;
    };
    PlanHost ph;
    PlanDev &p = ph.dev;
    std::vector<int32_t> lists[kNumClasses];
    std::vector<HubChunk> chunks;
    std::vector<int32_t> hub_start;
    std::vector<int32_t> hubs;
    std::vector<int32_t> zeros;
    if (g->sorted) {
        // first index in [a, b) with deg <= t (degrees non-increasing)
        auto first_le = [&](int32_t a, int32_t b, int64_t t) {
            while (a < b) {
                int32_t mid = a + (b - a) / 2;
                if (deg(mid) <= t) b = mid; else a = mid + 1;
            }
            return a;
        };
        // The degree band [dmin, dmax] is a contiguous range of rows.
        const int32_t band_beg = dmax == INT64_MAX ? start : first_le(start, end, dmax);
        const int32_t band_end = dmin <= 1 ? end : first_le(band_beg, end, dmin - 1);
        // Contiguous classes: class c is [first u with deg <= T_{c-1},
        // first u with deg <= T_c).
        const int64_t thr[kNumClasses] = {cfg.hub_threshold, cfg.block_threshold,
                                          cfg.wave_threshold, cfg.g16_threshold,
                                          cfg.g4_threshold, 0};
        int32_t lo = band_beg;
        for (int c = 0; c < kNumClasses; c++) {
            const int32_t a = first_le(lo, band_end, thr[c]);
            p.beg[c] = lo;
            p.cnt[c] = a - lo;
            lo = a;
        }
        p.zero_beg = lo;
        p.zero_cnt = dmin > 0 ? 0 : band_end - lo;
        for (int32_t u = p.beg[0]; u < p.beg[0] + p.cnt[0]; u++) hubs.push_back(u);
    } else {
        for (int32_t u = start; u < end; u++) {
            const int64_t d = deg(u);
            if (d > dmax || (d < dmin && !(d == 0 && dmin == 0))) continue;
            int c = cls_of(d);
            if (c < 0) {
                zeros.push_back(u);
                continue;
            }
            if (c == 0) hubs.push_back(u);
            else lists[c].push_back(u);
        }
        for (int c = 1; c < kNumClasses; c++) {
            p.beg[c] = 0;
            p.cnt[c] = (int32_t)lists[c].size();
        }
        p.zero_beg = 0;
        p.zero_cnt = (int32_t)zeros.size();
    }
    // Hub chunks.
    for (size_t h = 0; h < hubs.size(); h++) {
        int32_t u = hubs[h];
        int64_t b = g->h_off[u - g->row0], e = g->h_off[u - g->row0 + 1];
        hub_start.push_back((int32_t)chunks.size());
        for (int64_t x = b; x < e; x += cfg.hub_chunk)
            chunks.push_back(HubChunk{u, (int32_t)h, x, std::min(e, x + cfg.hub_chunk)});
    }
    hub_start.push_back((int32_t)chunks.size());
    p.beg[0] = 0;
    p.cnt[0] = (int32_t)chunks.size();
    p.num_hubs = (int32_t)hubs.size();
    p.unit_start[0] = 0;
    for (int c = 0; c < kNumClasses; c++) {
        int32_t units = (p.cnt[c] + kClassPerUnit[c] - 1) / kClassPerUnit[c];
        p.unit_start[c + 1] = p.unit_start[c] + units;
    }
    ph.total_units = p.unit_start[kNumClasses];
    // One device allocation: chunks | hub_start | lists...
    size_t bytes = chunks.size() * sizeof(HubChunk);
    size_t off_hub = bytes;
    bytes += hub_start.size() * sizeof(int32_t);
    size_t off_list[kNumClasses] = {0};
    for (int c = 1; c < kNumClasses; c++) {
        bytes = (bytes + 7) & ~size_t(7);
        off_list[c] = bytes;
        bytes += lists[c].size() * sizeof(int32_t);
    }
    bytes = (bytes + 7) & ~size_t(7);
    size_t off_zero = bytes;
    bytes += zeros.size() * sizeof(int32_t);
    bytes = std::max<size_t>(bytes, 8);
    HC_HIP(hipMalloc(&ph.dmem, bytes));
    char *base = static_cast<char *>(ph.dmem);
    if (!chunks.empty())
        HC_HIP(hipMemcpy(base, chunks.data(), chunks.size() * sizeof(HubChunk),
                         hipMemcpyHostToDevice));
    HC_HIP(hipMemcpy(base + off_hub, hub_start.data(),
                     hub_start.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    p.chunks = reinterpret_cast<const HubChunk *>(base);
    p.hub_chunk_start = reinterpret_cast<const int32_t *>(base + off_hub);
    p.list[0] = nullptr;
    for (int c = 1; c < kNumClasses; c++) {
        if (g->sorted) {
            p.list[c] = nullptr;
        } else {
            if (!lists[c].empty())
                HC_HIP(hipMemcpy(base + off_list[c], lists[c].data(),
                                 lists[c].size() * sizeof(int32_t),
                                 hipMemcpyHostToDevice));
            p.list[c] = reinterpret_cast<const int32_t *>(base + off_list[c]);
        }
    }
    if (g->sorted) {
        p.zero_list = nullptr;
    } else {
        if (!zeros.empty())
            HC_HIP(hipMemcpy(base + off_zero, zeros.data(),
                             zeros.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        p.zero_list = reinterpret_cast<const int32_t *>(base + off_zero);
    }
    auto res = g->plans.emplace(key, ph);
    *out = &res.first->second;
    return HC_OK;
}

static int upload_rows(const hc_host_graph_t *hg, int32_t start, int32_t end,
                       hc_graph_t *out) {
    HC_CHECK_ARG(hg && out, "NULL argument");
    HC_CHECK_ARG(hg->index, "graph index is NULL");
    HC_CHECK_ARG(hg->num_nodes >= 0, "num_nodes < 0");
    HC_CHECK_ARG(0 <= start && start <= end && end <= hg->num_nodes,
                 "row range [%d,%d) outside [0,%d)", start, end, hg->num_nodes);
    HC_CHECK_ARG(hg->num_edges == 0 || hg->neighbors,
                 "graph neighbors are NULL");
    HC_TRY(ensure_init());
    *out = nullptr;
    const int64_t *idx = hg->index;
    HC_CHECK_ARG(idx[0] == 0 && idx[hg->num_nodes] == hg->num_edges,
                 "index[0] must be 0 and index[N] must equal num_edges");
    const int32_t rows = end - start;
    const int64_t e0 = idx[start], e1 = idx[end];
    const int64_t m = e1 - e0;
    auto *g = new (std::nothrow) hc_graph_s();
    if (!g) return fail(HC_ERR_OOM, "host allocation failed");
    g->n = hg->num_nodes;
    g->row0 = start;
    g->rows = rows;
    g->m = m;
    g->weighted = hg->weights != nullptr;
    // 4-byte offsets while every offset plus an unrolled stride stays far
    // from 2^32; HC_FORCE_OFF64=1 selects 8-byte offsets for testing.
    const char *force64 = getenv("HC_FORCE_OFF64");
    g->off64 = m >= (int64_t(1) << 31) || (force64 && force64[0] == '1');
    g->h_off.resize(size_t(rows) + 1);
    bool sorted = true;
    int64_t maxd = 0;
    int32_t nz = 0;
    for (int32_t i = 0; i <= rows; i++) {
        g->h_off[i] = idx[start + i] - e0;
        if (i > 0) {
            int64_t d = g->h_off[i] - g->h_off[i - 1];
            if (d < 0) {
                delete g;
                return fail(HC_ERR_INVALID, "index not monotone at row %d",
                            start + i - 1);
            }
            if (i > 1 && d > g->h_off[i - 1] - g->h_off[i - 2]) sorted = false;
            maxd = std::max(maxd, d);
            if (d > 0) nz = i;
        }
    }
    g->sorted = sorted;
    g->max_degree = maxd;
    if (g->weighted && m > 0) {
        // Largest weight of an even sample of up to 2^16 edges (bucket width
        // of the batched sweep).
        const int32_t *wt = hg->weights + e0;
        const int64_t step = std::max<int64_t>(1, m >> 16);
        int32_t wmax = 0;
        for (int64_t i = 0; i < m; i += step) wmax = std::max(wmax, wt[i]);
        g->max_weight = wmax;
    }
    // For sorted rows, [row0, row0+nz) holds every non-zero-degree row.
    g->num_nonzero = sorted ? nz : rows;
    const int32_t *nb = hg->neighbors + e0;
    size_t ob = g->off64 ? 8 : 4;
    size_t off_bytes = ob * (size_t(rows) + 1);
    size_t adj_bytes = sizeof(int32_t) * size_t(std::max<int64_t>(m, 1));
    hipError_t e1r = hipMalloc(&g->d_off, off_bytes);
    hipError_t e2r = hipMalloc(reinterpret_cast<void **>(&g->d_adj), adj_bytes);
    hipError_t e3r = hipSuccess;
    if (g->weighted)
        e3r = hipMalloc(reinterpret_cast<void **>(&g->d_wt), adj_bytes);
    if (e1r != hipSuccess || e2r != hipSuccess || e3r != hipSuccess) {
        hipFree(g->d_off);
        hipFree(g->d_adj);
        hipFree(g->d_wt);
        delete g;
        return fail(HC_ERR_OOM, "device allocation of the CSR (%zu bytes) failed",
                    off_bytes + adj_bytes * (hg->weights ? 2 : 1));
    }
    g->bytes = off_bytes + adj_bytes * (g->weighted ? 2 : 1);
    // Offsets narrowed on the way; neighbour ids and weights checked while
    // they are packed into the pinned staging buffers (the reference reads
    // them unchecked).
    int rc = staged_offsets(g->d_off, g->h_off.data(), int64_t(rows) + 1, g->off64);
    int64_t bad = -1;
    if (rc == HC_OK && m > 0) {
        rc = staged_copy_checked(g->d_adj, nb, m, 0, hg->num_nodes, &bad);
        if (rc == HC_ERR_INVALID)
            rc = fail(HC_ERR_INVALID, "neighbour %lld = %d outside [0,%d)",
                      (long long)(e0 + bad), nb[bad], hg->num_nodes);
    }
    if (rc == HC_OK && g->weighted && m > 0) {
        const int32_t *w = hg->weights + e0;
        rc = staged_copy_checked(g->d_wt, w, m, 0, HC_INF_DIST, &bad);
        if (rc == HC_ERR_INVALID)
            rc = fail(HC_ERR_INVALID, "weight %lld = %d outside [0, 2^30)",
                      (long long)(e0 + bad), w[bad]);
    }
    if (rc != HC_OK) {
        hc_graph_free(g);
        return rc;
    }
    *out = g;
    return HC_OK;
}

}  // namespace hc

using namespace hc;

extern "C" {

int hc_init(int device) {
    DeviceState &s = device_state();
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(HC_ERR_NO_DEVICE, "no HIP device available (%s)",
                    e != hipSuccess ? hipGetErrorString(e) : "count 0");
    HC_CHECK_ARG(device >= 0 && device < count, "device %d of %d", device, count);
    if (s.ready && s.device == device) {
        HC_HIP(hipSetDevice(device));
        return HC_OK;
    }
    HC_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    HC_HIP(hipGetDeviceProperties(&prop, device));
    s.device = device;
    s.cus = prop.multiProcessorCount;
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            s.clock_khz = khz;
    }
    s.hbm = prop.totalGlobalMem;
    if (!s.stream) HC_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    s.ready = true;
    return HC_OK;
}

int hc_get_device_info(hc_device_info_t *out) {
    HC_CHECK_ARG(out, "out is NULL");
    HC_TRY(ensure_init());
    hipDeviceProp_t prop;
    HC_HIP(hipGetDeviceProperties(&prop, device_state().device));
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", prop.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
    if (out->name[0] == '\0') {
        // The marketing name comes from a table that some driver installs lack;
        // name the device by its ISA and CU count instead of leaving it blank.
        const char *colon = std::strchr(prop.gcnArchName, ':');
        int isa_len = colon ? (int)(colon - prop.gcnArchName) : (int)std::strlen(prop.gcnArchName);
        std::snprintf(out->name, sizeof(out->name), "AMD %.*s (%d CUs)", isa_len,
                      prop.gcnArchName, prop.multiProcessorCount);
    }
    out->compute_units = prop.multiProcessorCount;
    out->wavefront_size = prop.warpSize;
    out->hbm_bytes = prop.totalGlobalMem;
    out->lds_bytes_per_block = (int32_t)prop.sharedMemPerBlock;
    return HC_OK;
}

int hc_device_synchronize(void) {
    HC_TRY(ensure_init());
    HC_HIP(hipDeviceSynchronize());
    return HC_OK;
}

int hc_mem_get_info(size_t *free_bytes, size_t *total_bytes) {
    HC_CHECK_ARG(free_bytes && total_bytes, "NULL argument");
    HC_TRY(ensure_init());
    HC_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return HC_OK;
}

int hc_dmalloc(void **ptr, size_t bytes) {
    HC_CHECK_ARG(ptr, "ptr is NULL");
    HC_TRY(ensure_init());
    *ptr = nullptr;
    if (hipMalloc(ptr, std::max<size_t>(bytes, 4)) != hipSuccess)
        return fail(HC_ERR_OOM, "hipMalloc(%zu) failed", bytes);
    return HC_OK;
}

int hc_dfree(void *ptr) {
    if (!ptr) return HC_OK;
    HC_HIP(hipFree(ptr));
    return HC_OK;
}

int hc_hmalloc_pinned(void **ptr, size_t bytes) {
    HC_CHECK_ARG(ptr, "ptr is NULL");
    HC_TRY(ensure_init());
    if (hipHostMalloc(ptr, std::max<size_t>(bytes, 4), hipHostMallocDefault) != hipSuccess)
        return fail(HC_ERR_OOM, "hipHostMalloc(%zu) failed", bytes);
    return HC_OK;
}

int hc_hfree_pinned(void *ptr) {
    if (!ptr) return HC_OK;
    HC_HIP(hipHostFree(ptr));
    return HC_OK;
}

int hc_memcpy(void *dst, const void *src, size_t bytes, int kind,
              hc_stream_t stream) {
    HC_CHECK_ARG(kind >= 1 && kind <= 3, "kind must be 1, 2 or 3");
    if (bytes == 0) return HC_OK;
    HC_CHECK_ARG(dst && src, "NULL pointer");
    HC_TRY(ensure_init());
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                      : kind == 2 ? hipMemcpyDeviceToHost
                                  : hipMemcpyDeviceToDevice;
    hipStream_t s = pick_stream(stream);
    HC_HIP(hipMemcpyAsync(dst, src, bytes, k, s));
    HC_HIP(hipStreamSynchronize(s));
    return HC_OK;
}

__global__ void fill_i32_kernel(int32_t *p, int32_t v, size_t n) {
    size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int hc_memset_d32(void *dst, int32_t value, size_t count, hc_stream_t stream) {
    if (count == 0) return HC_OK;
    HC_CHECK_ARG(dst, "dst is NULL");
    HC_TRY(ensure_init());
    hipStream_t s = pick_stream(stream);
    int blocks = (int)std::min<size_t>((count + 255) / 256, size_t(grid_blocks()));
    fill_i32_kernel<<<blocks, 256, 0, s>>>(static_cast<int32_t *>(dst), value, count);
    HC_HIP(hipGetLastError());
    HC_HIP(hipStreamSynchronize(s));
    return HC_OK;
}

int hc_stream_create(hc_stream_t *stream) {
    HC_CHECK_ARG(stream, "stream is NULL");
    HC_TRY(ensure_init());
    hipStream_t s;
    HC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return HC_OK;
}

int hc_stream_destroy(hc_stream_t stream) {
    if (!stream) return HC_OK;
    HC_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return HC_OK;
}

int hc_stream_synchronize(hc_stream_t stream) {
    HC_TRY(ensure_init());
    HC_HIP(hipStreamSynchronize(pick_stream(stream)));
    return HC_OK;
}

int hc_graph_upload(const hc_host_graph_t *g, hc_graph_t *out) {
    HC_CHECK_ARG(g, "graph is NULL");
    return upload_rows(g, 0, g->num_nodes, out);
}

int hc_graph_upload_range(const hc_host_graph_t *g, int32_t start_id,
                          int32_t end_id, hc_graph_t *out) {
    HC_CHECK_ARG(g, "graph is NULL");
    return upload_rows(g, start_id, end_id, out);
}

int hc_graph_free(hc_graph_t g) {
    if (!g) return HC_OK;
    free_sssp_ws(g);
    free_bfs_ws(g);
    free_pr_ws(g);
    free_multi_ws(g);
    free_msbfs_ws(g);
    for (auto &kv : g->plans) hipFree(kv.second.dmem);
    if (!g->shared) {
        hipFree(g->d_off);
        hipFree(g->d_adj);
        hipFree(g->d_wt);
    }
    delete g;
    return HC_OK;
}

int hc_graph_share(hc_graph_t g, hc_graph_t *out) {
    HC_CHECK_ARG(g && out, "NULL argument");
    auto *h = new (std::nothrow) hc_graph_s();
    if (!h) return fail(HC_ERR_OOM, "host allocation failed");
    std::lock_guard<std::mutex> lk(g->mu);
    h->n = g->n;
    h->row0 = g->row0;
    h->rows = g->rows;
    h->m = g->m;
    h->weighted = g->weighted;
    h->sorted = g->sorted;
    h->off64 = g->off64;
    h->num_nonzero = g->num_nonzero;
    h->max_degree = g->max_degree;
    h->max_weight = g->max_weight;
    h->d_off = g->d_off;
    h->d_adj = g->d_adj;
    h->d_wt = g->d_wt;
    h->bytes = 0;
    h->h_off = g->h_off;
    h->shared = true;
    h->cbslot = g->cbslot;  // one column-blocked copy for the graph and its shares
    h->tune = g->tune;      // and one set of per-graph knobs
    *out = h;
    return HC_OK;
}

}  // extern "C"

namespace hc {
int set_knob(Tuning &t, const char *key, int64_t value) {
    HC_CHECK_ARG(key, "key is NULL");
    bool is_int = false;
    int *iref = nullptr;
    int64_t *p = knob_ref(t, key, &is_int, &iref);
    if (!p) return fail(HC_ERR_NOT_FOUND, "unknown tuning key '%s'", key);
    if (is_int) {
        const bool stage = !strcmp(key, "pull.stage") || !strcmp(key, "bu.stage") ||
                           !strcmp(key, "sssp.stage") || !strcmp(key, "push.stage");
        const bool flag = !strcmp(key, "profile.kernel_events") || !strcmp(key, "multi.graph") ||
                          !strcmp(key, "msbfs.on");
        if (!strcmp(key, "sssp.dist16"))
            HC_CHECK_ARG(value >= 0 && value <= 2, "sssp.dist16 must be 0, 1 or 2");
        if (!strcmp(key, "sssp.push_dedup"))
            HC_CHECK_ARG(value >= 0 && value <= 2, "sssp.push_dedup must be 0, 1 or 2");
        if (!strcmp(key, "bu.first"))
            HC_CHECK_ARG(value >= 0 && value <= 2, "bu.first must be 0, 1 or 2");
        if (!strcmp(key, "msbfs.col_shift")) {
            HC_CHECK_ARG(value == -1 || value == 0 || (value >= 6 && value <= 30),
                         "msbfs.col_shift must be -1 (automatic), 0 (off) or in [6, 30]");
            *iref = (int)value;
            return HC_OK;
        }
        if (!strcmp(key, "pull.col_shift")) {
            HC_CHECK_ARG(value == -1 || value == 0 || (value >= 6 && value <= 30),
                         "pull.col_shift must be -1 (automatic), 0 (off) or in [6, 30]");
            *iref = (int)value;
            return HC_OK;
        }
        if (!strcmp(key, "loop.batch"))
            HC_CHECK_ARG(value >= 1 && value <= 16, "loop.batch must be in [1, 16]");
        if (!strcmp(key, "multi.delta") || !strcmp(key, "sssp.delta")) {
            HC_CHECK_ARG(value >= -1 && value <= 1 << 30,
                         "%s must be -1 (automatic), 0 (off) or in [1, 2^30]", key);
            *iref = (int)value;
            return HC_OK;
        }
        if (!strcmp(key, "multi.block")) {
            HC_CHECK_ARG(value == 128 || value == 256 || value == 512 || value == 1024,
                         "multi.block must be 128, 256, 512 or 1024");
            *iref = (int)value;
            return HC_OK;
        }
        if (!strncmp(key, "many.", 5)) {
            if (!strcmp(key, "many.block"))
                HC_CHECK_ARG(value == 256 || value == 512 || value == 1024,
                             "many.block must be 256, 512 or 1024");
            if (!strcmp(key, "many.per_cu"))
                HC_CHECK_ARG(value >= 0 && value <= 32, "many.per_cu must be in [0, 32]");
            if (!strcmp(key, "many.heavy_cap"))
                HC_CHECK_ARG(value >= 0 && value <= kManyHeavyCapMax,
                             "many.heavy_cap must be in [0, %d]", kManyHeavyCapMax);
            if (!strcmp(key, "many.lds_nodes"))
                HC_CHECK_ARG(value >= 0 && value <= kManyLdsNodesMax,
                             "many.lds_nodes must be in [0, %d]", kManyLdsNodesMax);
            HC_CHECK_ARG(value >= 0 && value <= 1 << 20, "%s out of range", key);
            *iref = (int)value;
            return HC_OK;
        }
        if (!strcmp(key, "push.chunk"))
            HC_CHECK_ARG(value >= kMinChunkEdges, "push.chunk must be >= %d", (int)kMinChunkEdges);
        if (!strcmp(key, "bu.remaining"))
            HC_CHECK_ARG(value == 0 || value == 1 || value == 4 || value == 16,
                         "bu.remaining must be 0, 1, 4 or 16");
        const bool zero_ok = stage || flag || !strcmp(key, "bu.remaining") ||
                             !strcmp(key, "multi.push_div") || !strcmp(key, "sssp.filter") ||
                             !strcmp(key, "pull.col_shift") ||
                             !strcmp(key, "bu.first") || !strcmp(key, "sssp.alpha") ||
                             !strcmp(key, "sssp.dist16") || !strcmp(key, "sssp.push_dedup") ||
                             !strcmp(key, "bu.stage_min");
        HC_CHECK_ARG(value >= (zero_ok ? 0 : 1) && value <= (flag ? 1 : 1 << 20),
                     "%s out of range", key);
        if (stage)
            HC_CHECK_ARG(value * 4 <= kMaxStageBytes, "%s above %d entries", key,
                         kMaxStageBytes / 4);
        *iref = (int)value;
    } else {
        HC_CHECK_ARG(value >= 0, "%s must be >= 0", key);
        if (!strcmp(key, "pull.hub_chunk") || !strcmp(key, "multi.hub_chunk"))
            HC_CHECK_ARG(value >= 256, "%s must be >= 256", key);
        *p = value;
    }
    return HC_OK;
}

int check_knob(const char *key, int64_t value) {
    Tuning probe = global_tuning();
    return set_knob(probe, key, value);
}
}  // namespace hc

extern "C" {

int hc_set_tuning(const char *key, int64_t value) {
    return set_knob(global_tuning(), key, value);
}

int hc_get_tuning(const char *key, int64_t *value) {
    HC_CHECK_ARG(key && value, "NULL argument");
    bool is_int = false;
    int *iref = nullptr;
    int64_t *p = knob_ref(global_tuning(), key, &is_int, &iref);
    if (!p) return fail(HC_ERR_NOT_FOUND, "unknown tuning key '%s'", key);
    *value = is_int ? *iref : *p;
    return HC_OK;
}

int hc_reset_tuning(void) {
    global_tuning() = Tuning();
    return HC_OK;
}

int hc_graph_set_tuning(hc_graph_t g, const char *key, int64_t value) {
    HC_CHECK_ARG(g && key, "NULL argument");
    HC_TRY(check_knob(key, value));  // a rejected value is not recorded
    GraphTuning &gt = *g->tune;
    std::lock_guard<std::mutex> lk(gt.mu);
    for (auto &kv : gt.over)
        if (kv.first == key) {
            kv.second = value;
            return HC_OK;
        }
    gt.over.emplace_back(key, value);
    return HC_OK;
}

int hc_graph_get_tuning(hc_graph_t g, const char *key, int64_t *value) {
    HC_CHECK_ARG(g && key && value, "NULL argument");
    {
        GraphTuning &gt = *g->tune;
        std::lock_guard<std::mutex> lk(gt.mu);
        for (const auto &kv : gt.over)
            if (kv.first == key) {
                *value = kv.second;
                return HC_OK;
            }
    }
    return hc_get_tuning(key, value);
}

int hc_graph_reset_tuning(hc_graph_t g) {
    HC_CHECK_ARG(g, "NULL graph");
    GraphTuning &gt = *g->tune;
    std::lock_guard<std::mutex> lk(gt.mu);
    gt.over.clear();
    return HC_OK;
}

int hc_graph_get_info(hc_graph_t g, hc_graph_info_t *out) {
    HC_CHECK_ARG(g && out, "NULL argument");
    out->num_nodes = g->n;
    out->num_edges = g->m;
    out->start_id = g->row0;
    out->end_id = g->row0 + g->rows;
    out->degree_sorted = g->sorted ? 1 : 0;
    out->weighted = g->weighted ? 1 : 0;
    out->offset_bytes = g->off64 ? 8 : 4;
    out->num_nonzero = g->num_nonzero;
    out->max_degree = g->max_degree;
    out->device_bytes = g->bytes;
    return HC_OK;
}

}  // extern "C"
