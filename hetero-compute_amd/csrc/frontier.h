// Device-resident frontier and epoch-control state shared by the
// device-driven epoch loops (SSSP push/pull and BFS top-down/bottom-up).
//
// An epoch k reads the frontier of parity k&1 and writes parity (k+1)&1.
// Appended vertices are routed by degree into classes that map to different
// lane counts when the frontier is expanded:
//   chunk: deg > kHubDeg      -> one workgroup per kChunkEdges-edge chunk
//   q[0] : deg > kBigDeg      -> one workgroup per vertex
//   q[1] : deg > kMidDeg      -> one wavefront per vertex
//   q[2] : otherwise          -> kSmallLanes lanes per vertex
// The last workgroup of every step launch runs the controller, so the host
// can enqueue many steps without reading anything back.
#ifndef HC_FRONTIER_H
#define HC_FRONTIER_H

#include "device_common.h"

namespace hc {

constexpr int64_t kHubDeg = 16384;
constexpr int64_t kChunkEdges = 4096;
constexpr int64_t kBigDeg = 1024;
constexpr int64_t kMidDeg = 48;
constexpr int kSmallLanes = 8;
constexpr int kSmallPerUnit = kBlock / kSmallLanes;

enum StepMode : int32_t { kPush = 0, kPull = 1 };

struct StepCtl {
    uint32_t cnt[2][5];            // [parity][q0,q1,q2,chunks,hub vertices]
    unsigned long long edges[2];   // sum of degrees appended per parity
    uint32_t blocks_done;
    int32_t done;
    int32_t mode;
    int32_t epoch;
    int32_t max_epochs;
    int32_t direction;             // 0 auto, 1 pull only, 2 push only
    int32_t alpha;
    int32_t beta;
    long long m;                   // edges of the graph
    long long edges_to_check;      // BFS direction-optimizing state
    long long scout;
    int32_t num_nodes;             // BFS pull phase: current frontier size
    int32_t prev_nodes;
    int32_t n;                     // vertices
    int32_t push_epochs;
    int32_t pull_epochs;
    int32_t pad;
    unsigned long long reached;
    unsigned long long alg_bytes;  // algorithmic bytes moved by the epochs
    unsigned long long scanned;    // edges examined by the running epoch
    int32_t plan_rows;             // rows swept by a pull epoch
    int32_t pad2;
};

struct FrontierBufs {
    int32_t *q[2][3];
    HubChunk *ch[2];
    uint32_t *bm[3];
    int32_t words;                 // bitmap words
};

// Workspace layout for one graph (SSSP and BFS share the shape).
struct FrontierWs {
    FrontierBufs fb{};
    StepCtl *ctl = nullptr;
    StepCtl *h_ctl = nullptr;      // pinned mirror for polling
    void *dmem = nullptr;
    int64_t chunk_cap = 0;
};

int frontier_ws_alloc(hc_graph_s *g, FrontierWs *ws);
void frontier_ws_free(FrontierWs *ws);

template <typename OffT>
__device__ __forceinline__ void frontier_append(const GraphView<OffT> &g,
                                                const FrontierBufs &fb,
                                                StepCtl *ctl, int nxt,
                                                bool pred, int32_t v,
                                                unsigned long long &edges_acc) {
    int cls = -1;
    OffT b = 0, e = 0;
    if (pred) {
        b = g.rbeg(v);
        e = g.rend(v);
        int64_t d = int64_t(e - b);
        edges_acc += (unsigned long long)d;
        cls = d > kHubDeg ? 3 : (d > kBigDeg ? 0 : (d > kMidDeg ? 1 : 2));
    }
    wave_append(fb.q[nxt][0], &ctl->cnt[nxt][0], cls == 0, v);
    wave_append(fb.q[nxt][1], &ctl->cnt[nxt][1], cls == 1, v);
    wave_append(fb.q[nxt][2], &ctl->cnt[nxt][2], cls == 2, v);
    if (cls == 3) {
        int64_t d = int64_t(e - b);
        uint32_t nch = (uint32_t)((d + kChunkEdges - 1) / kChunkEdges);
        uint32_t base = atomicAdd(&ctl->cnt[nxt][3], nch);
        atomicAdd(&ctl->cnt[nxt][4], 1u);
        for (uint32_t j = 0; j < nch; j++) {
            int64_t cb = int64_t(b) + int64_t(j) * kChunkEdges;
            int64_t ce = cb + kChunkEdges < int64_t(e) ? cb + kChunkEdges : int64_t(e);
            fb.ch[nxt][base + j] = HubChunk{v, 0, cb, ce};
        }
    }
}

// Sum of a per-thread value over the workgroup; result valid in thread 0.
__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
    __shared__ unsigned long long s_part[kWavesPerBlock];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    __syncthreads();
    if (lane_id() == 0) s_part[threadIdx.x / kWave] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kWavesPerBlock; i++) t += s_part[i];
    return t;
}

}  // namespace hc

#endif
