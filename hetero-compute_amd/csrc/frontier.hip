// Frontier workspace allocation (device queues, chunk lists, bitmaps and the
// epoch controller block) shared by the SSSP and BFS epoch loops.
#include <algorithm>

#include "frontier.h"

namespace hc {

int frontier_ws_alloc(hc_graph_s *g, FrontierWs *ws) {
    const int64_t n = g->n;
    const int32_t words = (int32_t)((n + 31) / 32);
    // Chunk capacity: every hub vertex can be appended at most once per
    // epoch, so the sum of its chunk counts bounds one parity's list.
    int64_t cap = 1;
    for (int32_t i = 0; i < g->rows; i++) {
        int64_t d = g->h_off[i + 1] - g->h_off[i];
        if (d > kHubDeg) cap += (d + kChunkEdges - 1) / kChunkEdges;
    }
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t qbytes = al(sizeof(int32_t) * std::max<int64_t>(n, 1));
    size_t cbytes = al(sizeof(HubChunk) * cap);
    size_t bbytes = al(sizeof(uint32_t) * std::max(words, 1));
    size_t total = 6 * qbytes + 2 * cbytes + 3 * bbytes + al(sizeof(StepCtl));
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
    ws->ctl = reinterpret_cast<StepCtl *>(base + o);
    ws->fb.words = words;
    ws->dmem = base;
    ws->chunk_cap = cap;
    if (hipHostMalloc(reinterpret_cast<void **>(&ws->h_ctl), sizeof(StepCtl),
                      hipHostMallocDefault) != hipSuccess) {
        hipFree(base);
        ws->dmem = nullptr;
        return fail(HC_ERR_OOM, "pinned control block allocation failed");
    }
    return HC_OK;
}

void frontier_ws_free(FrontierWs *ws) {
    if (ws->dmem) hipFree(ws->dmem);
    if (ws->h_ctl) hipHostFree(ws->h_ctl);
    ws->dmem = nullptr;
    ws->h_ctl = nullptr;
}

}  // namespace hc
