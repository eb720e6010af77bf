// C++ drop-in for the reference's bitmap (src/bitmap.cuh:13-190): 32-bit
// words, bit (n & 31) of word n >> 5, a plain struct with free functions in
// namespace Bitmap.  Header-only host code over the C ABI of libhcamd.so.
//
// Host bitmaps (constructor / destructor / set_bit / set_bit_atomic /
// get_bit / reset / copy) keep the reference's behaviour; use_cuda = true
// allocates pinned host memory through hc_hmalloc_pinned.  The cu_cpu_*
// family keeps the reference's "struct on the host, buffer in HBM" objects
// (hc_dmalloc, hc_memset_d32, hc_memcpy), so their buffers can be passed to
// hc_bfs_pull_epoch; cu_constructor copies such a struct into device memory.
// Device-side bit operations (cu_set_bit_atomic) are the library's kernels'
// business and have no host counterpart here.
#ifndef HC_CPP_BITMAP_H
#define HC_CPP_BITMAP_H

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "../hc_capi.h"
#include "graph.h"

namespace Bitmap {

using data_t = int;
const std::size_t data_size = sizeof(data_t) * 8;

inline std::size_t data_offset(std::size_t n) { return n / data_size; }
inline std::size_t bit_offset(std::size_t n) { return n & (data_size - 1); }

struct Bitmap {
    data_t *buffer;
    bool using_cuda;
    std::size_t size;  // words
};

inline void reset(Bitmap *const bitmap) {
    std::fill(bitmap->buffer, bitmap->buffer + bitmap->size, 0);
}

inline Bitmap *constructor(const std::size_t num_elements, const bool use_cuda = false) {
    Bitmap *b = new Bitmap;
    b->size = (num_elements + data_size - 1) / data_size;
    b->using_cuda = use_cuda;
    if (use_cuda) {
        void *p = nullptr;
        hc::check_or_exit(hc_hmalloc_pinned(&p, std::max<std::size_t>(b->size, 1) * sizeof(data_t)),
                          "Bitmap::constructor");
        b->buffer = static_cast<data_t *>(p);
    } else {
        b->buffer = new data_t[std::max<std::size_t>(b->size, 1)];
    }
    reset(b);
    return b;
}

inline void destructor(Bitmap **const bitmap) {
    if ((*bitmap)->using_cuda)
        hc_hfree_pinned((*bitmap)->buffer);
    else
        delete[] (*bitmap)->buffer;
    delete *bitmap;
    *bitmap = nullptr;
}

inline void set_bit_atomic(Bitmap *const bitmap, const std::size_t idx) {
    const data_t bit = static_cast<data_t>(1u << bit_offset(idx));
    __atomic_fetch_or(&bitmap->buffer[data_offset(idx)], bit, __ATOMIC_RELAXED);
}

inline void set_bit(Bitmap *const bitmap, const std::size_t idx) {
    bitmap->buffer[data_offset(idx)] |= static_cast<data_t>(1u << bit_offset(idx));
}

inline bool get_bit(const Bitmap *const bitmap, const std::size_t idx) {
    return (static_cast<uint32_t>(bitmap->buffer[data_offset(idx)]) >> bit_offset(idx)) & 1u;
}

inline bool copy(const Bitmap *const from, Bitmap *const to) {
    if (from->size != to->size) return false;
    std::copy(from->buffer, from->buffer + from->size, to->buffer);
    return true;
}

// ---- bitmaps whose buffer lives in HBM ----

inline void cu_cpu_reset(Bitmap *const cu_bitmap) {
    hc::check_or_exit(hc_memset_d32(cu_bitmap->buffer, 0, cu_bitmap->size, nullptr),
                      "Bitmap::cu_cpu_reset");
}

inline Bitmap *cu_cpu_constructor(const std::size_t num_elements) {
    hc::check_or_exit(hc_init(0), "hc_init");
    Bitmap *b = new Bitmap;
    b->size = (num_elements + data_size - 1) / data_size;
    b->using_cuda = true;
    void *p = nullptr;
    hc::check_or_exit(hc_dmalloc(&p, std::max<std::size_t>(b->size, 1) * sizeof(data_t)),
                      "Bitmap::cu_cpu_constructor");
    b->buffer = static_cast<data_t *>(p);
    cu_cpu_reset(b);
    return b;
}

inline void cu_cpu_destructor(Bitmap **const cu_bitmap) {
    hc_dfree((*cu_bitmap)->buffer);
    delete *cu_bitmap;
    *cu_bitmap = nullptr;
}

// Sets one bit of an HBM bitmap from the host (word read, or, write back).
inline void cu_cpu_set_bit(Bitmap *const cu_bitmap, const std::size_t idx) {
    data_t w = 0;
    data_t *at = cu_bitmap->buffer + data_offset(idx);
    hc::check_or_exit(hc_memcpy(&w, at, sizeof(data_t), 2, nullptr), "Bitmap::cu_cpu_set_bit");
    w |= static_cast<data_t>(1u << bit_offset(idx));
    hc::check_or_exit(hc_memcpy(at, &w, sizeof(data_t), 1, nullptr), "Bitmap::cu_cpu_set_bit");
}

// Copy of a cu_cpu_constructor struct in device memory.
inline Bitmap *cu_constructor(const Bitmap *const bitmap) {
    void *p = nullptr;
    hc::check_or_exit(hc_dmalloc(&p, sizeof(Bitmap)), "Bitmap::cu_constructor");
    hc::check_or_exit(hc_memcpy(p, bitmap, sizeof(Bitmap), 1, nullptr), "Bitmap::cu_constructor");
    return static_cast<Bitmap *>(p);
}

inline void cu_destructor(Bitmap **const cu_bitmap) {
    hc_dfree(*cu_bitmap);
    *cu_bitmap = nullptr;
}

}  // namespace Bitmap

#endif  // HC_CPP_BITMAP_H
