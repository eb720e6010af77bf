// Internal helpers shared by the host (g++) and device (hipcc) halves of the
// library: error state, status codes, small checked-call macros.
#ifndef HC_INTERNAL_H
#define HC_INTERNAL_H

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/hc_capi.h"

namespace hc {

// Records a formatted message for hc_last_error() on this thread and returns
// the status code so call sites can `return fail(...)`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Upper bound on the device bytes the workspaces of one graph handle can
// hold once every single-source and batched algorithm has run on it (SSSP
// and BFS frontier queues, bitmaps and hub-chunk lists, BFS depth and
// first-neighbour arrays, the 16-bit distance mirror, batched-sweep flags,
// plans), for a graph of n vertices and m directed edges.
int64_t share_workspace_bound(int64_t n, int64_t m);
// Upper bound on the device bytes of the copies one graph (with all its
// shares) builds on first use: its column-blocked CSR copies.
int64_t graph_copy_bound(int64_t n, int64_t m, bool weighted);

// Mix of (vertex, value) for result checksums (splitmix64's finaliser,
// Vigna): the checksum of an array is the wrapping sum of mix over its
// entries, so any order of summation gives the same value.
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint64_t checksum_mix(uint64_t v, int32_t x) {
    uint64_t z = (v << 32) ^ (uint64_t)(uint32_t)x;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// HC_OK when key = value would be accepted by hc_set_tuning (nothing set).
int check_knob(const char *key, int64_t value);

// Host -> device copies through the library's pinned staging ring (packed by
// all host threads, sent asynchronously, overlapped; one upload at a time).
// staged_copy_checked: count int32 values, each checked against [lo, hi);
// returns HC_ERR_INVALID with *bad_index = the first offending index (no
// message set), else -1.  staged_offsets: count int64 offsets narrowed to
// uint32 (wide = false) or widened to uint64.
int staged_copy_checked(void *dst, const int32_t *src, int64_t count, int32_t lo, int32_t hi,
                        int64_t *bad_index);
int staged_offsets(void *dst, const int64_t *off, int64_t count, bool wide);

}  // namespace hc

#define HC_CHECK_ARG(cond, ...)                                   \
    do {                                                          \
        if (!(cond)) return hc::fail(HC_ERR_INVALID, __VA_ARGS__); \
    } while (0)

#define HC_TRY(expr)               \
    do {                           \
        int hc_rc_ = (expr);       \
        if (hc_rc_ != HC_OK) return hc_rc_; \
    } while (0)

#endif
