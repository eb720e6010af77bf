// C++ drop-in for the reference's host frontier queues, SlidingWindow and
// LocalWindow (src/window.h:19-143): a single buffer in which the current
// frontier is [start, end) and the next one is appended after it, and a
// per-thread staging buffer that reserves space in the shared buffer with one
// atomic add per flush.  Header-only, host code.
//
// Same names, members and semantics; as in the reference, push_back into the
// shared window is not thread-safe and no call checks for overflow (size the
// window for every vertex that can ever be appended, e.g. num_nodes).
#ifndef HC_CPP_WINDOW_H
#define HC_CPP_WINDOW_H

#include <algorithm>
#include <cstddef>
#include <memory>

template <typename T>
class LocalWindow;

template <typename T>
class SlidingWindow {
public:
    explicit SlidingWindow(size_t buffer_size) : buf(new T[buffer_size]) { reset(); }
    SlidingWindow(const SlidingWindow &) = delete;
    SlidingWindow &operator=(const SlidingWindow &) = delete;

    // Appends to the next window.
    void push_back(T val) { buf[next_end++] = val; }
    // The next window becomes the current one.
    void slide_window() {
        current_start = current_end;
        current_end = next_end;
    }
    void reset() { current_start = current_end = next_end = 0; }

    using iterator = T *;
    bool empty() const { return current_start == current_end; }
    iterator begin() const { return buf.get() + current_start; }
    iterator end() const { return buf.get() + current_end; }
    size_t size() const { return current_end - current_start; }

    friend class LocalWindow<T>;

private:
    std::unique_ptr<T[]> buf;
    size_t current_start = 0;  // current window [current_start, current_end)
    size_t current_end = 0;
    size_t next_end = 0;       // next window [current_end, next_end)
};

template <typename T>
class LocalWindow {
public:
    explicit LocalWindow(SlidingWindow<T> &parent_, size_t local_buffer_size = 16384)
        : parent(parent_), buf(new T[local_buffer_size]), cap(local_buffer_size) {}
    LocalWindow(const LocalWindow &) = delete;
    LocalWindow &operator=(const LocalWindow &) = delete;

    // Stages a value; a full buffer is flushed first.
    void push_back(T val) {
        if (count == cap) flush();
        buf[count++] = val;
    }
    // Appends the staged values to the parent's next window (safe to call
    // from many threads at once: the space is reserved atomically).
    void flush() {
        if (count == 0) return;
        const size_t at = __atomic_fetch_add(&parent.next_end, count, __ATOMIC_RELAXED);
        std::copy(buf.get(), buf.get() + count, parent.buf.get() + at);
        count = 0;
    }

private:
    SlidingWindow<T> &parent;
    std::unique_ptr<T[]> buf;
    size_t count = 0;
    const size_t cap;
};

#endif  // HC_CPP_WINDOW_H
