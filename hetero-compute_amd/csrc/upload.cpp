// Host -> HBM uploads through a pinned staging ring (compiled with g++).
//
// copy_subgraph_to_device (src/graph.cuh:250-286) copies the CSR arrays with
// plain cudaMemcpy from pageable memory: the driver stages every byte through
// its own bounce buffer, one call at a time.  Here the source is packed by all
// host threads into one of four pinned 32 MiB buffers (narrowing the int64
// offsets, or checking neighbour ids and weights, on the way), and each buffer
// is sent with an asynchronous copy on a stream of its own while the threads
// pack the next one: packing and the DMA overlap, and the DMA reads pinned
// pages at full link rate.
#include <hip/hip_runtime.h>
#include <omp.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <mutex>

#include "hc_internal.h"

namespace hc {

namespace {

constexpr int kStageBufs = 4;
constexpr size_t kStageBytes = size_t(32) << 20;

struct StagingRing {
    std::mutex mu;  // one upload at a time through the ring
    bool ready = false;
    char *buf[kStageBufs] = {};
    hipEvent_t sent[kStageBufs] = {};
    bool pending[kStageBufs] = {};
    hipStream_t stream = nullptr;

    int init() {
        if (ready) return HC_OK;
        for (int i = 0; i < kStageBufs; i++) {
            if (hipHostMalloc(reinterpret_cast<void **>(&buf[i]), kStageBytes,
                              hipHostMallocDefault) != hipSuccess)
                return fail(HC_ERR_OOM, "pinned staging buffer of %zu bytes failed",
                            kStageBytes);
            if (hipEventCreateWithFlags(&sent[i], hipEventDisableTiming) != hipSuccess)
                return fail(HC_ERR_HIP, "staging event creation failed");
        }
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
            return fail(HC_ERR_HIP, "staging stream creation failed");
        ready = true;
        return HC_OK;
    }
};

StagingRing &ring() {
    // Held until the process ends (the pinned buffers outlive every upload).
    static StagingRing *r = new StagingRing();
    return *r;
}

}  // namespace

int staged_upload(void *dst, size_t bytes,
                  const std::function<int(char *, size_t, size_t)> &fill) {
    if (bytes == 0) return HC_OK;
    StagingRing &r = ring();
    std::lock_guard<std::mutex> lk(r.mu);
    HC_TRY(r.init());
    int rc = HC_OK;
    int slot = 0;
    for (size_t off = 0; off < bytes && rc == HC_OK; off += kStageBytes) {
        const size_t len = std::min(kStageBytes, bytes - off);
        if (r.pending[slot]) {
            if (hipEventSynchronize(r.sent[slot]) != hipSuccess) {
                rc = fail(HC_ERR_HIP, "staged copy failed");
                break;
            }
            r.pending[slot] = false;
        }
        rc = fill(r.buf[slot], off, len);
        if (rc != HC_OK) break;
        if (hipMemcpyAsync(static_cast<char *>(dst) + off, r.buf[slot], len,
                           hipMemcpyHostToDevice, r.stream) != hipSuccess ||
            hipEventRecord(r.sent[slot], r.stream) != hipSuccess) {
            rc = fail(HC_ERR_HIP, "staged copy of %zu bytes failed", len);
            break;
        }
        r.pending[slot] = true;
        slot = (slot + 1) % kStageBufs;
    }
    // Every buffer is idle again before the next upload (or an error return).
    if (hipStreamSynchronize(r.stream) != hipSuccess && rc == HC_OK)
        rc = fail(HC_ERR_HIP, "staged upload failed");
    for (bool &p : r.pending) p = false;
    return rc;
}

int staged_copy_checked(void *dst, const int32_t *src, int64_t count, int32_t lo, int32_t hi,
                        int64_t *bad_index) {
    *bad_index = -1;
    return staged_upload(dst, size_t(count) * sizeof(int32_t),
                         [&](char *pinned, size_t off, size_t len) -> int {
                             const int64_t b = int64_t(off / sizeof(int32_t));
                             const int64_t n = int64_t(len / sizeof(int32_t));
                             int32_t *out = reinterpret_cast<int32_t *>(pinned);
                             int64_t first_bad = INT64_MAX;
#pragma omp parallel for schedule(static) reduction(min : first_bad)
                             for (int64_t i = 0; i < n; i++) {
                                 const int32_t v = src[b + i];
                                 out[i] = v;
                                 if (v < lo || v >= hi) first_bad = std::min(first_bad, b + i);
                             }
                             if (first_bad != INT64_MAX) {
                                 *bad_index = first_bad;
                                 return HC_ERR_INVALID;
                             }
                             return HC_OK;
                         });
}

int staged_offsets(void *dst, const int64_t *off, int64_t count, bool wide) {
    const size_t ob = wide ? 8 : 4;
    return staged_upload(dst, size_t(count) * ob, [&](char *pinned, size_t o, size_t len) -> int {
        const int64_t b = int64_t(o / ob), n = int64_t(len / ob);
        if (wide) {
            uint64_t *out = reinterpret_cast<uint64_t *>(pinned);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; i++) out[i] = uint64_t(off[b + i]);
        } else {
            uint32_t *out = reinterpret_cast<uint32_t *>(pinned);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; i++) out[i] = uint32_t(off[b + i]);
        }
        return HC_OK;
    });
}

}  // namespace hc
