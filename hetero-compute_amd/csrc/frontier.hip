// Frontier workspace allocation (device queues, chunk lists, bitmaps,
// per-workgroup partials and the control block) and the host side of the
// device-driven epoch loop, shared by SSSP and BFS.
#include <algorithm>
#include <cstring>

#include "frontier.h"

namespace hc {

int frontier_ws_alloc(hc_graph_s *g, FrontierWs *ws) {
    const int64_t n = g->n;
    const int32_t words = (int32_t)((n + 31) / 32);
    const int32_t grid = grid_blocks();
    // Chunk capacity: every hub vertex can be in one parity's frontier at
    // most once, so the sum of its chunk counts bounds one parity's list.
    int64_t cap = 1;
    for (int32_t i = 0; i < g->rows; i++) {
        int64_t d = g->h_off[i + 1] - g->h_off[i];
        if (d > kHubDeg) cap += (d + kMinChunkEdges - 1) / kMinChunkEdges;
    }
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t qbytes = al(sizeof(int32_t) * std::max<int64_t>(n, 1));
    size_t cbytes = al(sizeof(HubChunk) * cap);
    size_t bbytes = al(sizeof(uint32_t) * std::max(words, 1));
    size_t pbytes = al(sizeof(unsigned long long) * kSumWords * kSumSets);
    size_t nbytes = al(sizeof(uint32_t) * 2 * 5);
    size_t total =
        6 * qbytes + 2 * cbytes + 4 * bbytes + pbytes + nbytes + al(2 * sizeof(StepCtl));
    char *base = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(&base), total) != hipSuccess)
        return fail(HC_ERR_OOM, "frontier workspace of %zu bytes failed", total);
    size_t o = 0;
    for (int p = 0; p < 2; p++)
        for (int c = 0; c < 3; c++) {
            ws->fb.q[p][c] = reinterpret_cast<int32_t *>(base + o);
            o += qbytes;
        }
    for (int p = 0; p < 2; p++) {
        ws->fb.ch[p] = reinterpret_cast<HubChunk *>(base + o);
        o += cbytes;
    }
    for (int p = 0; p < 3; p++) {
        ws->fb.bm[p] = reinterpret_cast<uint32_t *>(base + o);
        o += bbytes;
    }
    ws->fb.far = reinterpret_cast<uint32_t *>(base + o);
    o += bbytes;
    ws->fb.part.acc = reinterpret_cast<unsigned long long *>(base + o);
    o += pbytes;
    ws->fb.cnt = reinterpret_cast<uint32_t (*)[5]>(base + o);
    o += nbytes;

    ws->ctl = reinterpret_cast<StepCtl *>(base + o);
    ws->fb.words = words;
    ws->dmem = base;
    ws->chunk_cap = cap;
    ws->grid = grid;
    if (hipHostMalloc(reinterpret_cast<void **>(&ws->h_ctl), sizeof(StepCtl),
                      hipHostMallocDefault) != hipSuccess) {
        hipFree(base);
        ws->dmem = nullptr;
        return fail(HC_ERR_OOM, "pinned control block allocation failed");
    }
    if (hipHostGetDevicePointer(reinterpret_cast<void **>(&ws->fb.host_done), &ws->h_ctl->done,
                                0) != hipSuccess) {
        frontier_ws_free(ws);
        return fail(HC_ERR_HIP, "pinned control block is not device-visible");
    }
    return HC_OK;
}

int64_t share_workspace_bound(int64_t n, int64_t m) {
    const int64_t words = (n + 31) / 32;
    // frontier_ws_alloc: six queues, two hub-chunk lists (a chunk per
    // kMinChunkEdges edges of every hub row, plus one per hub, bounded by
    // m / kMinChunkEdges + m / kHubDeg + 1), four bitmaps, partials, control.
    const int64_t chunks = m / kMinChunkEdges + m / kHubDeg + 1;
    const int64_t frontier = 6 * 4 * n + 2 * (int64_t)sizeof(HubChunk) * chunks +
                             4 * 4 * words + (int64_t)sizeof(unsigned long long) *
                             kSumWords * kSumSets + 64 * 1024;
    const int64_t sssp = frontier + 2 * n;                    // + 16-bit mirror
    const int64_t bfs = frontier + 8 * n + 4 * (3 * (int64_t)grid_blocks() + 1);  // + depth, first
    const int64_t multi = 3 * 4 * ((n + 3) / 4) + 4096;        // change flags
    // Bit-parallel batched BFS: seen and three frontier masks (8 bytes per
    // vertex each), three vertex lists, counters.
    const int64_t msbfs = 4 * 8 * n + 3 * 4 * n + 4096;
    // Plans: id lists (graphs that are not degree-sorted) for the SSSP,
    // bottom-up and batched sweeps and the column-blocked sweeps (heavy and
    // light rows, the batched BFS's), and their hub-chunk descriptors.
    const int64_t plans = 6 * (4 * n + (int64_t)sizeof(HubChunk) * chunks + 4096);
    return sssp + bfs + multi + msbfs + plans;
}

int64_t graph_copy_bound(int64_t n, int64_t m, bool weighted) {
    // Column-blocked copies (at most two block sizes: the SSSP sweeps' and
    // the batched BFS's): ids (and weights), at most 32 blocks of offsets
    // of up to 8 bytes, and the epoch API's row bitmap.
    const int64_t one = 4 * m * (weighted ? 2 : 1) + 32 * 8 * (n + 1) + 4 * ((n + 31) / 32) + 4096;
    return 2 * one;
}

void frontier_ws_free(FrontierWs *ws) {
    if (ws->dmem) hipFree(ws->dmem);
    if (ws->h_ctl) hipHostFree(ws->h_ctl);
    ws->dmem = nullptr;
    ws->h_ctl = nullptr;
}

int read_final_ctl(const StepCtl *d, StepCtl *out, hipStream_t s) {
    StepCtl h[2];
    HC_HIP(hipMemcpyAsync(h, d, 2 * sizeof(StepCtl), hipMemcpyDeviceToHost, s));
    HC_HIP(hipStreamSynchronize(s));
    *out = h[0];
    if (h[1].epoch > h[0].epoch)
        std::memcpy(out, &h[1], sizeof(unsigned long long) * kCtlHeadWords);
    return HC_OK;
}

int run_epoch_loop(FrontierWs *ws, hipStream_t s, long long cap,
                   int (*launch_epoch)(void *, long long), void *arg, float *ms,
                   long long *launches, StepCtl *final_ctl) {
    hipEvent_t ev0, ev1, evb[2];
    HC_HIP(hipEventCreate(&ev0));
    HC_HIP(hipEventCreate(&ev1));
    // The batch events only pace the host: no system-scope fence (an L2
    // writeback per batch); the done word is a system-scope store of its own.
    const unsigned pace = hipEventDisableTiming | hipEventDisableSystemFence;
    HC_HIP(hipEventCreateWithFlags(&evb[0], pace));
    HC_HIP(hipEventCreateWithFlags(&evb[1], pace));
    StepCtl *h = ws->h_ctl;
    h->done = 0;
    HC_HIP(hipEventRecord(ev0, s));
    // Batches of two epochs: the host reads the done flag after batch b-1 has
    // finished while batch b is queued, so at most ~4 epochs are enqueued past
    // the last one (they return at once) and the device never waits on a poll.
    long long epochs = 0;
    int rc = HC_OK;
    bool done = false;
    int b = 0;
    while (!done && epochs < cap) {
        const int batch = tuning().loop_batch;
        for (int i = 0; i < batch; i++) {
            rc = launch_epoch(arg, epochs);
            if (rc != HC_OK) break;
            epochs++;
        }
        if (rc != HC_OK) break;
        // The controller sets h->done itself when the loop ends.
        hipEventRecord(evb[b & 1], s);
        if (b > 0) {
            hipEventSynchronize(evb[(b - 1) & 1]);
            done = __atomic_load_n(&h->done, __ATOMIC_ACQUIRE) != 0;
        }
        b++;
    }
    HC_HIP(hipEventRecord(ev1, s));
    HC_HIP(hipEventSynchronize(ev1));
    HC_HIP(hipEventElapsedTime(ms, ev0, ev1));
    rc = rc != HC_OK ? rc : read_final_ctl(ws->ctl, final_ctl, s);
    hipEventDestroy(ev0);
    hipEventDestroy(ev1);
    hipEventDestroy(evb[0]);
    hipEventDestroy(evb[1]);
    *launches = epochs;
    return rc;
}

}  // namespace hc

using namespace hc;

extern "C" int hc_get_epoch_trace(hc_graph_t g, int algorithm,
                                  hc_epoch_trace_t *out) {
    HC_CHECK_ARG(g && out, "NULL argument");
    HC_CHECK_ARG(algorithm == 0 || algorithm == 1, "algorithm must be 0 (SSSP) or 1 (BFS)");
    std::lock_guard<std::mutex> lk(g->mu);
    // Both workspaces start with their FrontierWs.
    void *ws = algorithm == 0 ? g->ws_sssp : g->ws_bfs;
    std::memset(out, 0, sizeof(*out));
    if (!ws) return HC_OK;
    StepCtl c;
    HC_HIP(hipDeviceSynchronize());
    HC_TRY(read_final_ctl(static_cast<FrontierWs *>(ws)->ctl, &c, nullptr));
    out->count = std::min(c.epoch, kTraceLen);
    const double khz = device_state().clock_khz;
    for (int i = 0; i < out->count; i++) {
        out->mode[i] = c.tr_mode[i];
        out->frontier[i] = c.tr_n[i];
        out->frontier_edges[i] = c.tr_e[i];
        out->scanned[i] = c.tr_scan[i];
        const long long t0 = i == 0 ? c.t_start : c.tr_t[i - 1];
        out->us[i] = khz > 0 ? double(c.tr_t[i] - t0) * 1e3 / khz : 0.0;
    }
    return HC_OK;
}
