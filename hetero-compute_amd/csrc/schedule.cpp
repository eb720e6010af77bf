// Segment schedules: kernel and device names, the .skd text format, and
// schedule objects (compiled with g++).
//
//  * Names: the reference's to_string() strings (src/kernels/cpu/sssp.cuh:
//    221-231, src/kernels/gpu/sssp.cuh:322-333, src/kernels/cpu/pr.cuh:
//    226-236, src/kernels/gpu/pr.cuh:359-371, src/kernels/cpu/bfs.cuh:480-490,
//    src/kernels/gpu/bfs.cuh:331-343; devices src/devices.h:48-57) plus the
//    degree-bucketed GPU sweeps and the two devices of this library.
//  * Parser: load_sssp_schedule (src/schedule.h:151-226) generalised to the
//    three algorithms, with kernel lookup by prefix (get_kernel_id,
//    src/schedule.h:133-142), and errors instead of the reference's
//    unchecked std::stoi / exit.
//  * Writer: the layout of write_schedule (util/scheduler/sched.py:263-286).
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hc_sched.h"
#include "hc_internal.h"

struct hc_schedule_s {
    std::vector<hc_schedule_block_t> blocks;
    int32_t num_segments = 0;
};

namespace {

struct KernelName {
    const char *name;
    int32_t device_type;
};

// Index = hc_kernel_id.  Lookup by prefix walks this table in order, so a
// name that is a prefix of another must come after it; none is.
const KernelName kKernels[HC_K_COUNT] = {
    {"SSSP CPU one-to-one", HC_DEV_CPU}, {"SSSP GPU one-to-one", HC_DEV_GPU},
    {"SSSP GPU warp-red", HC_DEV_GPU},   {"SSSP GPU block-red", HC_DEV_GPU},
    {"SSSP GPU bucketed", HC_DEV_GPU},   {"PR CPU one-to-one", HC_DEV_CPU},
    {"PR GPU one-to-one", HC_DEV_GPU},   {"PR GPU warp-red", HC_DEV_GPU},
    {"PR GPU block-red", HC_DEV_GPU},    {"PR GPU bucketed", HC_DEV_GPU},
    {"BFS CPU pull", HC_DEV_CPU},        {"BFS GPU one-to-one", HC_DEV_GPU},
    {"BFS GPU warp", HC_DEV_GPU},        {"BFS GPU bucketed", HC_DEV_GPU},
    {"BFS CPU push by node", HC_DEV_CPU}, {"BFS CPU push by edge", HC_DEV_CPU},
};

const KernelName kDevices[HC_DEVICE_COUNT] = {
    {"Intel i7-9700K", HC_DEV_CPU},
    {"Intel Xeon E5-2686", HC_DEV_CPU},
    {"NVIDIA Quadro RTX 4000", HC_DEV_GPU},
    {"NVIDIA Tesla K80", HC_DEV_GPU},
    {"NVIDIA Tesla M60", HC_DEV_GPU},
    {"AMD Instinct MI355X", HC_DEV_GPU},
    {"Host CPU", HC_DEV_CPU},
};

// Strips a trailing '\r' and surrounding blanks (files written on other
// systems; the reference's std::getline keeps them and then fails to match).
std::string trim(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && (s[b] == ' ' || s[b] == '\t')) b++;
    while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r' ||
                     s[e - 1] == '\n'))
        e--;
    return s.substr(b, e - b);
}

int device_lookup(const std::string &line) {
    for (int d = 0; d < HC_DEVICE_COUNT; d++)
        if (line == kDevices[d].name) return d;
    return -1;
}

int kernel_lookup(const std::string &line) {
    for (int k = 0; k < HC_K_COUNT; k++) {
        size_t len = std::strlen(kKernels[k].name);
        if (line.compare(0, len, kKernels[k].name) == 0 &&
            (line.size() == len || line[len] == ' '))
            return k;
    }
    return -1;
}

// Parses a non-negative int32 token; false on junk, overflow or trailing text.
bool parse_int(const std::string &tok, int32_t *out) {
    if (tok.empty()) return false;
    errno = 0;
    char *end = nullptr;
    long v = std::strtol(tok.c_str(), &end, 10);
    if (errno || *end != '\0' || v < 0 || v > INT32_MAX) return false;
    *out = (int32_t)v;
    return true;
}

std::vector<std::string> split(const std::string &s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

int validate_blocks(const hc_schedule_block_t *b, int32_t n, int32_t segs) {
    for (int32_t i = 0; i < n; i++) {
        HC_CHECK_ARG(b[i].device >= 0 && b[i].device < HC_DEVICE_COUNT,
                     "block %d: unknown device %d", i, b[i].device);
        HC_CHECK_ARG(b[i].kernel >= 0 && b[i].kernel < HC_K_COUNT,
                     "block %d: unknown kernel %d", i, b[i].kernel);
        HC_CHECK_ARG(kKernels[b[i].kernel].device_type ==
                         kDevices[b[i].device].device_type,
                     "block %d: kernel \"%s\" cannot run on \"%s\"", i,
                     kKernels[b[i].kernel].name, kDevices[b[i].device].name);
        HC_CHECK_ARG(b[i].device_id >= 0, "block %d: negative device id", i);
        HC_CHECK_ARG(0 <= b[i].start_segment &&
                         b[i].start_segment <= b[i].end_segment &&
                         b[i].end_segment < segs,
                     "block %d: segments [%d,%d] outside [0,%d)", i,
                     b[i].start_segment, b[i].end_segment, segs);
        HC_CHECK_ARG(b[i].blocks >= 0 && b[i].threads >= 0 &&
                         b[i].threads <= 1024,
                     "block %d: launch shape %d x %d out of range", i,
                     b[i].blocks, b[i].threads);
    }
    return HC_OK;
}

}  // namespace

extern "C" {

int hc_kernel_from_string(const char *name, int32_t *kernel) {
    HC_CHECK_ARG(name && kernel, "NULL argument");
    int k = kernel_lookup(trim(name));
    if (k < 0) return hc::fail(HC_ERR_NOT_FOUND, "unknown kernel \"%s\"", name);
    *kernel = k;
    return HC_OK;
}

const char *hc_kernel_to_string(int32_t kernel) {
    if (kernel < 0 || kernel >= HC_K_COUNT) return "undefined kernel";
    return kKernels[kernel].name;
}

int hc_kernel_device_type(int32_t kernel, int32_t *device_type) {
    HC_CHECK_ARG(device_type, "NULL argument");
    HC_CHECK_ARG(kernel >= 0 && kernel < HC_K_COUNT, "unknown kernel %d", kernel);
    *device_type = kKernels[kernel].device_type;
    return HC_OK;
}

int hc_device_from_string(const char *name, int32_t *device) {
    HC_CHECK_ARG(name && device, "NULL argument");
    int d = device_lookup(trim(name));
    if (d < 0) return hc::fail(HC_ERR_NOT_FOUND, "unknown device \"%s\"", name);
    *device = d;
    return HC_OK;
}

const char *hc_device_to_string(int32_t device) {
    if (device < 0 || device >= HC_DEVICE_COUNT) return "undefined device";
    return kDevices[device].name;
}

int hc_device_type_of(int32_t device, int32_t *device_type) {
    HC_CHECK_ARG(device_type, "NULL argument");
    HC_CHECK_ARG(device >= 0 && device < HC_DEVICE_COUNT, "unknown device %d", device);
    *device_type = kDevices[device].device_type;
    return HC_OK;
}

int hc_schedule_create(const hc_schedule_block_t *blocks, int32_t num_blocks,
                       int32_t num_segments, hc_schedule_t *out) {
    HC_CHECK_ARG(out, "NULL argument");
    HC_CHECK_ARG(num_blocks > 0 && blocks, "a schedule needs at least one block");
    HC_CHECK_ARG(num_segments > 0, "num_segments must be positive");
    HC_TRY(validate_blocks(blocks, num_blocks, num_segments));
    auto *s = new (std::nothrow) hc_schedule_s;
    if (!s) return hc::fail(HC_ERR_OOM, "out of host memory");
    s->blocks.assign(blocks, blocks + num_blocks);
    s->num_segments = num_segments;
    *out = s;
    return HC_OK;
}

int hc_schedule_parse(const char *text, hc_schedule_t *out) {
    HC_CHECK_ARG(text && out, "NULL argument");
    std::istringstream is(text);
    std::vector<hc_schedule_block_t> blocks;
    std::map<int, int> instances;  // device -> appearances so far
    int cur_device = -1, cur_id = 0, max_seg = -1;
    std::string raw;
    int lineno = 0;
    while (std::getline(is, raw)) {
        lineno++;
        std::string line = trim(raw);
        if (line.empty()) continue;
        int d = device_lookup(line);
        if (d >= 0) {
            cur_device = d;
            cur_id = instances[d]++;
            continue;
        }
        HC_CHECK_ARG(cur_device >= 0,
                     "line %d: \"%s\" is neither a device name nor after one",
                     lineno, line.c_str());
        std::vector<std::string> segs = split(line);
        hc_schedule_block_t b{};
        HC_CHECK_ARG(segs.size() == 2 && parse_int(segs[0], &b.start_segment) &&
                         parse_int(segs[1], &b.end_segment) &&
                         b.start_segment <= b.end_segment,
                     "line %d: expected \"<start_segment> <end_segment>\", got \"%s\"",
                     lineno, line.c_str());
        std::string kline;
        do {
            HC_CHECK_ARG(std::getline(is, raw), "line %d: segment range without a kernel line",
                         lineno);
            lineno++;
            kline = trim(raw);
        } while (kline.empty());
        int k = kernel_lookup(kline);
        HC_CHECK_ARG(k >= 0, "line %d: unknown kernel \"%s\"", lineno, kline.c_str());
        HC_CHECK_ARG(kKernels[k].device_type == kDevices[cur_device].device_type,
                     "line %d: kernel \"%s\" cannot run on \"%s\"", lineno,
                     kKernels[k].name, kDevices[cur_device].name);
        std::vector<std::string> extra = split(kline.substr(std::strlen(kKernels[k].name)));
        if (kKernels[k].device_type == HC_DEV_GPU) {
            // The reference requires "<blocks> <threads>" after a GPU kernel
            // (src/schedule.h:199-206); here it is optional (0 = default).
            HC_CHECK_ARG(extra.empty() ||
                             (extra.size() == 2 && parse_int(extra[0], &b.blocks) &&
                              parse_int(extra[1], &b.threads) && b.threads <= 1024),
                         "line %d: expected \"%s <blocks> <threads>\", got \"%s\"",
                         lineno, kKernels[k].name, kline.c_str());
        } else {
            HC_CHECK_ARG(extra.empty(), "line %d: CPU kernel \"%s\" takes no launch shape",
                         lineno, kKernels[k].name);
        }
        b.device = cur_device;
        b.device_id = cur_id;
        b.kernel = k;
        blocks.push_back(b);
        max_seg = std::max(max_seg, b.end_segment);
    }
    HC_CHECK_ARG(!blocks.empty(), "schedule holds no blocks");
    return hc_schedule_create(blocks.data(), (int32_t)blocks.size(), max_seg + 1, out);
}

int hc_schedule_load(const char *path, hc_schedule_t *out) {
    HC_CHECK_ARG(path && out, "NULL argument");
    std::ifstream ifs(path, std::ios::in | std::ios::binary);
    if (!ifs) return hc::fail(HC_ERR_IO, "cannot open schedule \"%s\"", path);
    std::stringstream ss;
    ss << ifs.rdbuf();
    if (ifs.bad()) return hc::fail(HC_ERR_IO, "cannot read schedule \"%s\"", path);
    return hc_schedule_parse(ss.str().c_str(), out);
}

int hc_schedule_info(hc_schedule_t s, int32_t *num_blocks, int32_t *num_segments) {
    HC_CHECK_ARG(s, "NULL schedule");
    if (num_blocks) *num_blocks = (int32_t)s->blocks.size();
    if (num_segments) *num_segments = s->num_segments;
    return HC_OK;
}

int hc_schedule_get_block(hc_schedule_t s, int32_t i, hc_schedule_block_t *out) {
    HC_CHECK_ARG(s && out, "NULL argument");
    HC_CHECK_ARG(i >= 0 && i < (int32_t)s->blocks.size(), "block %d outside [0,%d)", i,
                 (int32_t)s->blocks.size());
    *out = s->blocks[i];
    return HC_OK;
}

int hc_schedule_to_text(hc_schedule_t s, char *buf, size_t cap, size_t *needed) {
    HC_CHECK_ARG(s, "NULL schedule");
    // One section per device instance, in order of first appearance, as
    // write_schedule emits one per DeviceSchedule (sched.py:279-286).
    std::vector<std::pair<int, int>> inst;
    for (const hc_schedule_block_t &b : s->blocks) {
        std::pair<int, int> k{b.device, b.device_id};
        if (std::find(inst.begin(), inst.end(), k) == inst.end()) inst.push_back(k);
    }
    std::string text;
    for (const auto &k : inst) {
        text += kDevices[k.first].name;
        text += '\n';
        for (const hc_schedule_block_t &b : s->blocks) {
            if (b.device != k.first || b.device_id != k.second) continue;
            text += std::to_string(b.start_segment) + ' ' + std::to_string(b.end_segment) + '\n';
            text += kKernels[b.kernel].name;
            if (kKernels[b.kernel].device_type == HC_DEV_GPU)
                text += ' ' + std::to_string(b.blocks) + ' ' + std::to_string(b.threads);
            text += '\n';
        }
    }
    if (needed) *needed = text.size() + 1;
    if (!buf) return HC_OK;
    HC_CHECK_ARG(cap >= text.size() + 1, "buffer of %zu bytes, %zu needed", cap,
                 text.size() + 1);
    std::memcpy(buf, text.c_str(), text.size() + 1);
    return HC_OK;
}

int hc_schedule_free(hc_schedule_t s) {
    delete s;
    return HC_OK;
}

}  // extern "C"
