// Serving loop: resident and dynamic graphs, BFS / SSSP queries admitted by
// device-memory footprint and batched as they arrive (continuous batching).
// Host C++ (g++) over the library's C ABI; one worker thread and HIP stream
// per slot.  No reference counterpart: the reference runs one query per
// driver call (src/kernels/gpu/sssp.cuh:42-100, src/kernels/gpu/bfs.cuh:
// 21-95) and uploads the graph each time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hc_batch.h"
#include "../../include/hc_serve.h"
#include "hc_internal.h"

namespace {

constexpr int kWidth = HC_MULTI_WIDTH;

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Device result storage shared by the queries of one sweep.
struct Result {
    std::mutex mu;                // held while the device copy is read
    int32_t *p = nullptr;         // device copy (nullptr once spilled)
    size_t bytes = 0;
    int refs = 0;
    bool budgeted = true;         // bytes still counted in the budget
    bool spilling = false;
    std::vector<int32_t> host;    // host copy after a spill
};

struct Query {
    int32_t graph = 0, kind = 0, source = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int32_t status = HC_OK;
    int32_t batch = 0, slot = -1;
    bool done = false;
    std::string error;  // message of a failed run
    std::shared_ptr<Result> res;
    int32_t column = -1;  // column of a batched result, -1 for a single one
    bool has_sum = false;  // sum computed after the sweep (opts.checksum)
    uint64_t sum = 0;
};

// Arrival rate of one (graph, kind) class: moving average of the gaps
// between its submissions.
struct ClassRate {
    double last = -1.0;  // ms of the last submission
    double gap = -1.0;   // ms between submissions (-1: fewer than two yet)
};

struct GraphEntry {
    hc_host_graph_t host{};
    bool resident = false;
    hc_graph_t dev = nullptr;             // owner of the device arrays
    std::vector<hc_graph_t> per_slot;     // shares, one per slot
    std::vector<std::pair<std::string, int64_t>> knobs;  // applied at every upload
    // Device ms of one single-source query and of one 64-wide batch, per
    // kind, measured when a resident graph is added (0: not measured).
    double single_ms[2] = {0, 0};
    double batch_ms[2] = {0, 0};
    int64_t bytes = 0;                    // budgeted bytes while uploaded
    bool uploading = false;
    int in_use = 0;
    double last_use = 0;
};

}  // namespace

struct hc_server_s {
    hc_server_opts_t opts{};
    bool user_min[2] = {false, false};  // min_batch_bfs / _sssp given by the caller
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_mem;
    std::deque<int64_t> queue;
    std::map<int64_t, Query> queries;   // stable iterators across inserts
    std::map<std::pair<int32_t, int32_t>, ClassRate> rates;  // by (graph, kind)
    std::deque<GraphEntry> graphs;      // stable references across inserts
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    std::multimap<size_t, int32_t *> pool;  // idle result buffers by size
    int64_t pooled = 0;                     // bytes held by the pool
    int64_t in_use = 0;                     // live results + dynamic graphs
    int64_t next_ticket = 1;
    int running = 0;
    int waiting_mem = 0;  // slots holding a batch until memory is released
    bool stop = false;
    hc_server_stats_t st{};
    // Fetch path (client threads, serialised).
    std::mutex fetch_mu;
    int32_t *fetch_buf = nullptr;
    size_t fetch_cap = 0;
    hipStream_t fetch_stream = nullptr;
};

namespace {

// Budget charge of a graph: per-slot workspaces at their upper bound, plus
// the CSR (device layout: 4- or 8-byte offsets, int32 ids and weights) for a
// dynamic graph.
int64_t graph_charge(const hc_server_s *s, const hc_host_graph_t *g, bool resident) {
    const int64_t n = g->num_nodes, m = g->num_edges;
    const int64_t off = (m < (int64_t(1) << 31) ? 4 : 8) * (n + 1);
    const int64_t csr = off + 4 * m * (g->weights ? 2 : 1);
    const int64_t ws = hc::share_workspace_bound(n, m) * s->opts.slots +
                       hc::graph_copy_bound(n, m, g->weights != nullptr);
    return resident ? ws : csr + ws;
}

int64_t node_bytes(const GraphEntry &g, int width) {
    return (int64_t)g.host.num_nodes * 4 * width;
}

void note_usage(hc_server_s *s) {
    s->st.bytes_in_use = s->in_use + s->pooled;
    s->st.bytes_peak = std::max(s->st.bytes_peak, s->st.bytes_in_use);
}

// Makes room for `need` more bytes under the budget: trims the buffer pool,
// then evicts idle dynamic graphs (not `keep`).  Called with mu held; may
// drop it around device frees.  Returns true when the bytes fit.
bool make_room(hc_server_s *s, std::unique_lock<std::mutex> &lk, int64_t need, int keep) {
    const int64_t budget = s->opts.memory_budget;
    while (s->in_use + s->pooled + need > budget && !s->pool.empty()) {
        auto it = s->pool.begin();
        int32_t *p = it->second;
        s->pooled -= (int64_t)it->first;
        s->pool.erase(it);
        lk.unlock();
        (void)hipFree(p);
        lk.lock();
    }
    while (s->in_use + s->pooled + need > budget) {
        int victim = -1;
        for (int i = 0; i < (int)s->graphs.size(); i++) {
            GraphEntry &g = s->graphs[i];
            if (i == keep || g.resident || !g.dev || g.in_use || g.uploading) continue;
            if (victim < 0 || g.last_use < s->graphs[victim].last_use) victim = i;
        }
        if (victim < 0) return false;
        GraphEntry &g = s->graphs[victim];
        std::vector<hc_graph_t> hs = g.per_slot;
        hc_graph_t owner = g.dev;
        g.per_slot.assign(g.per_slot.size(), nullptr);
        g.dev = nullptr;
        s->in_use -= g.bytes;
        s->st.evictions++;
        lk.unlock();
        for (hc_graph_t h : hs)
            if (h) hc_graph_free(h);
        hc_graph_free(owner);
        lk.lock();
    }
    return true;
}

int32_t *take_buffer(hc_server_s *s, size_t bytes) {
    auto it = s->pool.find(bytes);
    if (it == s->pool.end()) return nullptr;
    int32_t *p = it->second;
    s->pool.erase(it);
    s->pooled -= (int64_t)bytes;
    return p;
}

void finish(hc_server_s *s, const std::vector<int64_t> &batch, int status, double t_done) {
    const std::string msg = status != HC_OK ? hc_last_error() : "";
    for (int64_t t : batch) {
        Query &q = s->queries[t];
        q.status = status;
        q.error = msg;
        q.t_done = t_done;
        q.done = true;
        s->st.completed++;
        if (status != HC_OK) s->st.failed++;
    }
    s->cv_done.notify_all();
}

void release_result(hc_server_s *s, std::shared_ptr<Result> &r) {
    if (!r) return;
    if (--r->refs == 0 && !r->spilling) {
        if (r->budgeted) s->in_use -= (int64_t)r->bytes;
        r->budgeted = false;
        if (r->p) {
            s->pool.emplace(r->bytes, r->p);
            s->pooled += (int64_t)r->bytes;
        }
        r->p = nullptr;
        s->cv_mem.notify_all();
    }
    r.reset();
}

// Moves the largest finished, unreleased result in HBM to host memory and
// returns its buffer to the pool, so that admission in arrival order always
// progresses while clients hold results.  Called with mu held; drops it
// around the copy.  Returns false when nothing can be spilled.
bool spill_one(hc_server_s *s, std::unique_lock<std::mutex> &lk) {
    std::shared_ptr<Result> best;
    for (auto &kv : s->queries) {
        const Query &q = kv.second;
        if (!q.done || !q.res || !q.res->p || q.res->spilling) continue;
        if (!best || q.res->bytes > best->bytes) best = q.res;
    }
    if (!best) return false;
    best->spilling = true;
    lk.unlock();
    bool ok = true;
    {
        // The result's lock is held until its device copy has gone back to
        // the pool: a concurrent hc_server_fetch (which reads res->p under
        // this lock) sees either the device copy before any reuse or the
        // host copy.  Lock order: result, then server (fetch never holds
        // both).
        std::lock_guard<std::mutex> rl(best->mu);
        try {
            best->host.resize(best->bytes / sizeof(int32_t));
        } catch (...) {
            ok = false;
        }
        ok = ok && hipMemcpy(best->host.data(), best->p, best->bytes, hipMemcpyDeviceToHost) ==
                       hipSuccess;
        lk.lock();
        best->spilling = false;
        if (ok) {
            s->pool.emplace(best->bytes, best->p);
            s->pooled += (int64_t)best->bytes;
            best->p = nullptr;
            if (best->budgeted) s->in_use -= (int64_t)best->bytes;
            best->budgeted = false;
            s->st.spills++;
        } else {
            best->host.clear();
        }
    }
    if (best->refs == 0) {  // released while spilling
        std::shared_ptr<Result> r = best;
        r->refs = 1;
        release_result(s, r);
    }
    return ok;
}

// Chooses the next sweep (caller holds s->mu): classes (graph, kind) are
// taken in the arrival order of their oldest waiting query.  A class runs
// its oldest queries up to max_batch when at least its min_batch of them
// wait (class_min_batch), else its oldest query alone -- except that a class
// that batches at all and whose arrivals would gather 2 x min_batch queries
// (at most max_batch) within max_wait_ms of its oldest one's arrival waits
// (other classes run meanwhile) until max_batch wait or the oldest has
// waited max_wait_ms.  Returns false when every waiting class is gathering;
// *wake is then the ms at which the first of them stops.
// Waiting queries of a class from which a batched sweep beats running them
// one by one: the measured break-even count, or the fixed option; above
// max_batch the class never batches.
int class_min_batch(const hc_server_s *s, const GraphEntry &g, int kind) {
    const int fixed = kind == HC_QUERY_BFS ? s->opts.min_batch_bfs : s->opts.min_batch_sssp;
    if (s->user_min[kind] || !(g.single_ms[kind] > 0 && g.batch_ms[kind] > 0)) return fixed;
    const double k = std::ceil(g.batch_ms[kind] / g.single_ms[kind]);
    return (int)std::max(2.0, std::min(k, (double)s->opts.max_batch + 1));
}

bool pick_batch(hc_server_s *s, double now, std::vector<int64_t> &batch, double *wake) {
    *wake = 1e300;
    std::vector<std::pair<int32_t, int32_t>> seen;
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        const Query &head = s->queries[*it];
        const std::pair<int32_t, int32_t> cls{head.graph, head.kind};
        if (std::find(seen.begin(), seen.end(), cls) != seen.end()) continue;
        seen.push_back(cls);
        int waiting = 0;
        for (auto jt = it; jt != s->queue.end() && waiting < s->opts.max_batch; ++jt)
            if (s->queries[*jt].graph == cls.first && s->queries[*jt].kind == cls.second)
                waiting++;
        const int min_batch = class_min_batch(s, s->graphs[cls.first], cls.second);
        if (min_batch <= s->opts.max_batch && !s->stop && s->opts.max_wait_ms > 0 &&
            waiting < s->opts.max_batch) {
            const double gap = s->rates[cls].gap;
            const double deadline = head.t_submit + s->opts.max_wait_ms;
            const double target = std::min(2.0 * min_batch, (double)s->opts.max_batch);
            if (gap > 0 && now < deadline && waiting + (deadline - now) / gap >= target) {
                *wake = std::min(*wake, deadline);
                continue;  // gathering
            }
        }
        const int take = waiting >= min_batch && min_batch <= s->opts.max_batch
                             ? s->opts.max_batch
                             : 1;
        for (auto jt = it; jt != s->queue.end() && (int)batch.size() < take;) {
            const Query &q = s->queries[*jt];
            if (q.graph == cls.first && q.kind == cls.second) {
                batch.push_back(*jt);
                jt = s->queue.erase(jt);
            } else {
                ++jt;
            }
        }
        return true;
    }
    return false;
}

void worker(hc_server_s *s, int slot) {
    hipStream_t stream = s->streams[slot];
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // Admission is in arrival order: while one slot waits for memory no
        // other slot starts a younger batch, so results of older queries are
        // the only ones that can hold the budget.
        s->cv_work.wait(lk, [&] { return s->stop || (!s->queue.empty() && s->waiting_mem == 0); });
        if (s->queue.empty()) {
            if (s->stop) break;
            continue;
        }
        std::vector<int64_t> batch;
        double wake = 0;
        if (!pick_batch(s, now_ms(), batch, &wake)) {
            // Every waiting class is gathering: sleep until the first of them
            // is due, or until a submission (or shutdown) wakes this slot.
            s->cv_work.wait_for(lk, std::chrono::duration<double, std::milli>(
                                        std::max(0.05, wake - now_ms())));
            continue;
        }
        const int gid = s->queries[batch[0]].graph, kind = s->queries[batch[0]].kind;
        GraphEntry &g = s->graphs[gid];
        g.in_use++;
        // Admission: result memory (and the graph, if it is not in HBM).
        const int width = batch.size() > 1 ? kWidth : 1;
        const int64_t res_bytes = node_bytes(g, width);
        // Waits for released results when the budget is spent; a client
        // that holds results while waiting for later queries must release
        // them first.
        int status = HC_OK;
        for (;;) {
            while (g.uploading) s->cv_mem.wait(lk);
            const int64_t need = res_bytes + (g.dev ? 0 : g.bytes);
            if (need > s->opts.memory_budget) {
                status = hc::fail(HC_ERR_OOM, "query needs %lld bytes, budget %lld",
                                  (long long)need, (long long)s->opts.memory_budget);
                break;
            }
            const bool fits = make_room(s, lk, need, gid);
            if (g.uploading) continue;  // another slot began this graph's upload
            if (fits) break;
            if (spill_one(s, lk)) continue;
            if (s->stop && s->running == 0) {
                status = hc::fail(HC_ERR_BUSY, "server stopped with the memory budget spent");
                break;
            }
            s->waiting_mem++;
            s->cv_mem.wait(lk);
            s->waiting_mem--;
        }
        s->cv_work.notify_all();
        if (status != HC_OK) {
            g.in_use--;
            finish(s, batch, status, now_ms());
            continue;
        }
        s->in_use += res_bytes;
        bool upload = false;
        if (!g.dev) {
            s->in_use += g.bytes;
            g.uploading = true;
            upload = true;
        }
        note_usage(s);
        int32_t *buf = take_buffer(s, (size_t)res_bytes);
        std::vector<int32_t> sources;
        for (int64_t t : batch) sources.push_back(s->queries[t].source);
        s->running++;
        lk.unlock();

        // Device work, outside the lock.  A stale error of an earlier call
        // on this thread must not be charged to this batch.
        (void)hipGetLastError();
        if (upload) {
            hc_graph_t d = nullptr;
            status = hc_graph_upload(&g.host, &d);
            lk.lock();
            g.uploading = false;
            if (status == HC_OK) {
                g.dev = d;
                for (const auto &kv : g.knobs)  // checked when recorded
                    (void)hc_graph_set_tuning(d, kv.first.c_str(), kv.second);
                s->st.uploads++;
            } else {
                s->in_use -= g.bytes;
            }
            s->cv_mem.notify_all();
            lk.unlock();
        }
        hc_graph_t h = nullptr;
        if (status == HC_OK) {
            lk.lock();
            if (!g.per_slot[slot]) status = hc_graph_share(g.dev, &g.per_slot[slot]);
            h = g.per_slot[slot];
            lk.unlock();
        }
        if (status == HC_OK && !buf) {
            void *p = nullptr;
            status = hipMalloc(&p, (size_t)res_bytes) == hipSuccess
                         ? HC_OK
                         : hc::fail(HC_ERR_OOM, "result allocation of %lld bytes failed",
                                    (long long)res_bytes);
            buf = static_cast<int32_t *>(p);
        }
        const double t_start = now_ms();
        if (status == HC_OK) {
            hc_run_stats_t rs{};
            if (width == 1 && kind == HC_QUERY_BFS) {
                hc_bfs_opts_t o{15, 18, 0};
                status = hc_bfs(h, sources[0], buf, nullptr, &o, &rs, stream);
            } else if (width == 1) {
                status = hc_sssp(h, sources[0], buf, nullptr, &rs, stream);
            } else if (kind == HC_QUERY_BFS) {
                status = hc_bfs_multi(h, sources.data(), (int32_t)sources.size(), buf, 0, &rs,
                                      stream);
            } else {
                status = hc_sssp_multi(h, sources.data(), (int32_t)sources.size(), buf, 0, &rs,
                                       stream);
            }
            if (status == HC_OK && hipStreamSynchronize(stream) != hipSuccess)
                status = hc::fail(HC_ERR_HIP, "stream synchronisation failed");
        }
        const double t_done = now_ms();
        // Result checksums, on the slot's stream while the slot still
        // holds the sweep.  A failed checksum does not fail the queries:
        // hc_server_checksum then computes it on demand.
        uint64_t sums[kWidth] = {};
        bool summed = false;
        if (status == HC_OK && s->opts.checksum) {
            const int minus1 = kind == HC_QUERY_BFS ? 1 : 0;
            summed = (width == 1 ? hc_checksum_i32(buf, (int32_t)g.host.num_nodes, 1, 0, 0,
                                                   sums, stream)
                                 : hc_checksum_multi(buf, (int32_t)g.host.num_nodes,
                                                     (int32_t)batch.size(), minus1, sums,
                                                     stream)) == HC_OK;
            if (!summed) (void)hipGetLastError();
        }

        lk.lock();
        s->running--;
        g.in_use--;
        g.last_use = t_done;
        auto res = std::make_shared<Result>();
        res->p = buf;
        res->bytes = (size_t)res_bytes;
        res->refs = (int)batch.size();
        if (status != HC_OK) {
            res->refs = 1;
            std::shared_ptr<Result> r = res;
            release_result(s, r);
        }
        for (size_t i = 0; i < batch.size(); i++) {
            Query &q = s->queries[batch[i]];
            q.t_start = t_start;
            q.batch = (int32_t)batch.size();
            q.slot = slot;
            if (status == HC_OK) {
                q.res = res;
                q.column = width == 1 ? -1 : (int32_t)i;
                q.has_sum = summed;
                q.sum = sums[i];
            }
        }
        s->st.batches++;
        if (batch.size() > 1) s->st.batched_queries += (int64_t)batch.size();
        finish(s, batch, status, t_done);
        note_usage(s);
        s->cv_mem.notify_all();  // its results can now be spilled
    }
}

}  // namespace

extern "C" {

int hc_server_create(const hc_server_opts_t *opts, hc_server_t *out) {
    HC_CHECK_ARG(out, "NULL argument");
    HC_TRY(hc_init(0));
    auto *s = new (std::nothrow) hc_server_s();
    if (!s) return hc::fail(HC_ERR_OOM, "host allocation failed");
    s->opts.slots = opts && opts->slots > 0 ? opts->slots : 4;
    s->opts.max_batch = opts && opts->max_batch > 0 ? opts->max_batch : kWidth;
    s->opts.memory_budget = opts ? opts->memory_budget : 0;
    s->user_min[HC_QUERY_SSSP] = opts && opts->min_batch_sssp > 0;
    s->user_min[HC_QUERY_BFS] = opts && opts->min_batch_bfs > 0;
    s->opts.min_batch_sssp = opts && opts->min_batch_sssp > 0 ? opts->min_batch_sssp : 8;
    s->opts.min_batch_bfs = opts && opts->min_batch_bfs > 0 ? opts->min_batch_bfs : kWidth;
    s->opts.max_wait_ms = opts && opts->max_wait_ms != 0 ? opts->max_wait_ms : 500;
    s->opts.calibrate = opts ? opts->calibrate : 0;
    s->opts.checksum = opts ? opts->checksum : 0;
    if (s->opts.slots > 16 || s->opts.max_batch > kWidth) {
        delete s;
        return hc::fail(HC_ERR_INVALID, "slots must be <= 16 and max_batch <= %d", kWidth);
    }
    if (s->opts.memory_budget <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            delete s;
            return hc::fail(HC_ERR_HIP, "hipMemGetInfo failed");
        }
        s->opts.memory_budget = (int64_t)(free_b / 2);
    }
    s->streams.resize(s->opts.slots);
    for (int i = 0; i < s->opts.slots; i++)
        if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess) {
            for (int j = 0; j < i; j++) (void)hipStreamDestroy(s->streams[j]);
            delete s;
            return hc::fail(HC_ERR_HIP, "stream creation failed");
        }
    if (hipStreamCreateWithFlags(&s->fetch_stream, hipStreamNonBlocking) != hipSuccess) {
        for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
        delete s;
        return hc::fail(HC_ERR_HIP, "stream creation failed");
    }
    for (int i = 0; i < s->opts.slots; i++) s->workers.emplace_back(worker, s, i);
    *out = s;
    return HC_OK;
}

int hc_server_destroy(hc_server_t s) {
    if (!s) return HC_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_mem.notify_all();
    for (std::thread &t : s->workers) t.join();
    for (auto &kv : s->queries)
        if (kv.second.res && --kv.second.res->refs == 0 && kv.second.res->p)
            (void)hipFree(kv.second.res->p);
    for (auto &kv : s->pool) (void)hipFree(kv.second);
    for (GraphEntry &g : s->graphs) {
        for (hc_graph_t h : g.per_slot)
            if (h) hc_graph_free(h);
        if (g.dev) hc_graph_free(g.dev);
    }
    if (s->fetch_buf) (void)hipFree(s->fetch_buf);
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    (void)hipStreamDestroy(s->fetch_stream);
    delete s;
    return HC_OK;
}

}  // extern "C"

namespace {

// Device ms of single-source queries and of a 64-wide batch of each kind on
// a resident graph (e.dev), on a stream of its own (slots may be serving
// other graphs meanwhile): each run once to build plans and workspaces, then
// timed (singles averaged over three sources).  Sources are vertices with
// edges, spread over the id range.  Leaves the costs at 0 when the graph has
// no edges or the scratch matrix or stream cannot be allocated.
void calibrate_graph(hc_server_s *s, GraphEntry &e) {
    const int32_t n = e.host.num_nodes;
    std::vector<int32_t> src;
    const int64_t step = std::max<int64_t>(1, n / 257);
    for (int64_t k = 0; k < n && (int)src.size() < kWidth; k++) {
        const int32_t v = (int32_t)((k * step + k / 257) % n);
        if (e.host.index[v + 1] > e.host.index[v]) src.push_back(v);
    }
    if (src.empty()) return;
    for (size_t i = 0; (int)src.size() < kWidth; i++) {
        const int32_t v = src[i];  // repeat the ones found
        src.push_back(v);
    }
    void *buf = nullptr;
    if (hipMalloc(&buf, sizeof(int32_t) * (size_t)n * kWidth) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    int32_t *d = static_cast<int32_t *>(buf);
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(buf);
        return;
    }
    hc_run_stats_t rs{};
    for (int kind = 0; kind < 2; kind++) {
        if (kind == HC_QUERY_SSSP && !e.host.weights) continue;
        auto single = [&](int32_t v) {
            if (kind == HC_QUERY_BFS) {
                hc_bfs_opts_t o{15, 18, 0};
                return hc_bfs(e.dev, v, d, nullptr, &o, &rs, st);
            }
            return hc_sssp(e.dev, v, d, nullptr, &rs, st);
        };
        auto batch = [&]() {
            return kind == HC_QUERY_BFS ? hc_bfs_multi(e.dev, src.data(), kWidth, d, 0, &rs, st)
                                        : hc_sssp_multi(e.dev, src.data(), kWidth, d, 0, &rs, st);
        };
        if (single(src[0]) != HC_OK) continue;
        double sum = 0;
        int runs = 0;
        for (int i = 1; i <= 3 && i < (int)src.size(); i++)
            if (single(src[i * 7 % src.size()]) == HC_OK) {
                sum += rs.ms;
                runs++;
            }
        if (runs == 0 || batch() != HC_OK || batch() != HC_OK) continue;
        e.single_ms[kind] = sum / runs;
        e.batch_ms[kind] = rs.ms;
    }
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    hc::clear_error();
}

}  // namespace

extern "C" {

int hc_server_add_graph(hc_server_t s, const hc_host_graph_t *g, int32_t resident,
                        int32_t *graph_id) {
    HC_CHECK_ARG(s && g && graph_id, "NULL argument");
    HC_CHECK_ARG(g->num_nodes > 0 && g->index, "empty graph");
    GraphEntry e;
    e.host = *g;
    e.resident = resident != 0;
    e.per_slot.assign(s->opts.slots, nullptr);
    e.bytes = graph_charge(s, g, e.resident);
    if (e.resident) {
        std::lock_guard<std::mutex> lk(s->mu);
        HC_CHECK_ARG(s->in_use + s->pooled + e.bytes <= s->opts.memory_budget,
                     "the workspaces of a resident graph (%lld bytes) exceed the memory budget",
                     (long long)e.bytes);
    }
    if (e.resident) HC_TRY(hc_graph_upload(g, &e.dev));
    if (e.resident && s->opts.calibrate >= 0) calibrate_graph(s, e);
    std::lock_guard<std::mutex> lk(s->mu);
    if (e.resident) {
        // Resident graphs keep their per-slot workspaces for the server's
        // lifetime: charged once, here.
        s->in_use += e.bytes;
        s->st.uploads++;
        note_usage(s);
    }
    s->graphs.push_back(std::move(e));
    *graph_id = (int32_t)s->graphs.size() - 1;
    return HC_OK;
}

int hc_server_graph_costs(hc_server_t s, int32_t graph_id, double *single_ms,
                          double *batch_ms) {
    HC_CHECK_ARG(s && single_ms && batch_ms, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    HC_CHECK_ARG(graph_id >= 0 && graph_id < (int32_t)s->graphs.size(), "unknown graph %d",
                 graph_id);
    const GraphEntry &g = s->graphs[graph_id];
    for (int k = 0; k < 2; k++) {
        single_ms[k] = g.single_ms[k];
        batch_ms[k] = g.batch_ms[k];
    }
    return HC_OK;
}

int hc_server_set_graph_tuning(hc_server_t s, int32_t graph_id, const char *key,
                               int64_t value) {
    HC_CHECK_ARG(s && key, "NULL argument");
    HC_TRY(hc::check_knob(key, value));
    std::lock_guard<std::mutex> lk(s->mu);
    HC_CHECK_ARG(graph_id >= 0 && graph_id < (int32_t)s->graphs.size(), "unknown graph %d",
                 graph_id);
    GraphEntry &g = s->graphs[graph_id];
    bool found = false;
    for (auto &kv : g.knobs)
        if (kv.first == key) {
            kv.second = value;
            found = true;
        }
    if (!found) g.knobs.emplace_back(key, value);
    // The uploaded handle (and so every slot's share of it) takes it now.
    if (g.dev) HC_TRY(hc_graph_set_tuning(g.dev, key, value));
    return HC_OK;
}

int hc_server_graph_bytes(hc_server_t s, const hc_host_graph_t *g, int32_t resident,
                          int64_t *bytes) {
    HC_CHECK_ARG(s && g && bytes, "NULL argument");
    HC_CHECK_ARG(g->num_nodes > 0 && g->index, "empty graph");
    *bytes = graph_charge(s, g, resident != 0);
    return HC_OK;
}

int hc_server_submit(hc_server_t s, int32_t graph_id, int32_t kind, int32_t source,
                     int64_t *ticket) {
    HC_CHECK_ARG(s && ticket, "NULL argument");
    HC_CHECK_ARG(kind == HC_QUERY_BFS || kind == HC_QUERY_SSSP, "unknown query kind %d", kind);
    std::lock_guard<std::mutex> lk(s->mu);
    HC_CHECK_ARG(graph_id >= 0 && graph_id < (int)s->graphs.size(), "unknown graph %d", graph_id);
    const GraphEntry &g = s->graphs[graph_id];
    HC_CHECK_ARG(source >= 0 && source < g.host.num_nodes, "source %d outside [0,%d)", source,
                 g.host.num_nodes);
    HC_CHECK_ARG(kind == HC_QUERY_BFS || g.host.weights, "SSSP needs a weighted graph");
    HC_CHECK_ARG(!s->stop, "server is shutting down");
    const int64_t t = s->next_ticket++;
    Query q;
    q.graph = graph_id;
    q.kind = kind;
    q.source = source;
    q.t_submit = now_ms();
    ClassRate &r = s->rates[{graph_id, kind}];
    if (r.last >= 0) {
        const double g = q.t_submit - r.last;
        r.gap = r.gap < 0 ? g : 0.875 * r.gap + 0.125 * g;
    }
    r.last = q.t_submit;
    s->queries.emplace(t, std::move(q));
    s->queue.push_back(t);
    s->st.submitted++;
    *ticket = t;
    s->cv_work.notify_one();
    return HC_OK;
}

int hc_server_wait(hc_server_t s, int64_t ticket, hc_query_info_t *info) {
    HC_CHECK_ARG(s, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->queries.find(ticket);
    if (it == s->queries.end()) return hc::fail(HC_ERR_NOT_FOUND, "unknown ticket %lld", (long long)ticket);
    s->cv_done.wait(lk, [&] { return it->second.done; });
    const Query &q = it->second;
    // A failed query's message becomes this thread's hc_last_error().
    if (q.status != HC_OK) hc::fail(q.status, "%s", q.error.c_str());
    if (info) {
        const GraphEntry &g = s->graphs[q.graph];
        info->status = q.status;
        info->kind = q.kind;
        info->graph = q.graph;
        info->source = q.source;
        info->batch_size = q.batch;
        info->slot = q.slot;
        info->num_nodes = g.host.num_nodes;
        info->num_edges = g.host.num_edges;
        info->queue_ms = q.t_start - q.t_submit;
        info->run_ms = q.t_done - q.t_start;
        info->latency_ms = q.t_done - q.t_submit;
    }
    return HC_OK;
}

int hc_server_fetch(hc_server_t s, int64_t ticket, int32_t *host_out) {
    HC_CHECK_ARG(s && host_out, "NULL argument");
    std::shared_ptr<Result> res;
    int32_t column = -1, kind = 0;
    int64_t n = 0;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->queries.find(ticket);
        if (it == s->queries.end())
            return hc::fail(HC_ERR_NOT_FOUND, "unknown ticket %lld", (long long)ticket);
        s->cv_done.wait(lk, [&] { return it->second.done; });
        if (it->second.status != HC_OK)
            return hc::fail(it->second.status, "query %lld failed: %s", (long long)ticket,
                            it->second.error.c_str());
        res = it->second.res;
        column = it->second.column;
        kind = it->second.kind;
        n = s->graphs[it->second.graph].host.num_nodes;
    }
    std::lock_guard<std::mutex> rl(res->mu);
    if (!res->p) {  // spilled to host
        const bool minus1 = kind == HC_QUERY_BFS;
        for (int64_t v = 0; v < n; v++) {
            const int32_t x = column >= 0 ? res->host[(size_t)v * kWidth + column] : res->host[v];
            host_out[v] = (column >= 0 && minus1 && x == HC_INF_DIST) ? -1 : x;
        }
        return HC_OK;
    }
    std::lock_guard<std::mutex> fl(s->fetch_mu);
    const int32_t *src = res->p;
    if (column >= 0) {
        const size_t need = sizeof(int32_t) * (size_t)n;
        if (s->fetch_cap < need) {
            if (s->fetch_buf) (void)hipFree(s->fetch_buf);
            s->fetch_buf = nullptr;
            s->fetch_cap = 0;
            void *p = nullptr;
            if (hipMalloc(&p, need) != hipSuccess)
                return hc::fail(HC_ERR_OOM, "fetch buffer allocation failed");
            s->fetch_buf = static_cast<int32_t *>(p);
            s->fetch_cap = need;
        }
        HC_TRY(hc_multi_extract(res->p, (int32_t)n, column, kind == HC_QUERY_BFS ? 1 : 0,
                                s->fetch_buf, s->fetch_stream));
        src = s->fetch_buf;
    }
    if (hipMemcpyAsync(host_out, src, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost,
                       s->fetch_stream) != hipSuccess ||
        hipStreamSynchronize(s->fetch_stream) != hipSuccess)
        return hc::fail(HC_ERR_HIP, "result copy failed");
    return HC_OK;
}

int hc_server_checksum(hc_server_t s, int64_t ticket, uint64_t *sum) {
    HC_CHECK_ARG(s && sum, "NULL argument");
    std::shared_ptr<Result> res;
    int32_t column = -1, kind = 0;
    int64_t n = 0;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->queries.find(ticket);
        if (it == s->queries.end())
            return hc::fail(HC_ERR_NOT_FOUND, "unknown ticket %lld", (long long)ticket);
        s->cv_done.wait(lk, [&] { return it->second.done; });
        if (it->second.status != HC_OK)
            return hc::fail(it->second.status, "query %lld failed: %s", (long long)ticket,
                            it->second.error.c_str());
        if (it->second.has_sum) {
            *sum = it->second.sum;
            return HC_OK;
        }
        res = it->second.res;
        column = it->second.column;
        kind = it->second.kind;
        n = s->graphs[it->second.graph].host.num_nodes;
    }
    const bool minus1 = kind == HC_QUERY_BFS;
    std::lock_guard<std::mutex> rl(res->mu);
    if (!res->p) {  // spilled to host: the same sum on the CPU
        uint64_t h = 0;
        for (int64_t v = 0; v < n; v++) {
            int32_t x = column >= 0 ? res->host[(size_t)v * kWidth + column] : res->host[v];
            if (column >= 0 && minus1 && x == HC_INF_DIST) x = -1;
            h += hc::checksum_mix((uint64_t)v, x);
        }
        *sum = h;
        return HC_OK;
    }
    std::lock_guard<std::mutex> fl(s->fetch_mu);
    return hc_checksum_i32(res->p, (int32_t)n, column >= 0 ? kWidth : 1,
                           column >= 0 ? column : 0, column >= 0 && minus1 ? 1 : 0, sum,
                           s->fetch_stream);
}

int hc_server_release(hc_server_t s, int64_t ticket) {
    HC_CHECK_ARG(s, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->queries.find(ticket);
    if (it == s->queries.end()) return hc::fail(HC_ERR_NOT_FOUND, "unknown ticket %lld", (long long)ticket);
    s->cv_done.wait(lk, [&] { return it->second.done; });
    release_result(s, it->second.res);
    s->queries.erase(it);
    note_usage(s);
    return HC_OK;
}

int hc_server_get_stats(hc_server_t s, hc_server_stats_t *out) {
    HC_CHECK_ARG(s && out, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    return HC_OK;
}

}  // extern "C"
