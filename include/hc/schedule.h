// C++ drop-in for the reference's device names (src/devices.h:16-90) and
// schedule types (src/schedule.h:32-226): Device / DeviceType with
// to_string / get_device / get_device_type, ScheduleBlock and its SSSP CPU /
// GPU forms, the SSSPScheduleBlock variant, SSSPHeteroSchedule, their stream
// printers, get_kernel_id<EnumT> and load_sssp_schedule.  Header-only, over
// the validating .skd parser of libhcamd.so (hc_sched.h).
//
// Same names, members and printed layout.  Differences: the Device list adds
// this library's two devices (AMD Instinct MI355X, Host CPU); a GPU block's
// `kernel` is the SSSPGPU registry id (GPU kernels are not host function
// pointers here); load_sssp_schedule prints the parser's message and exits
// on a malformed or unreadable file, where the reference returns a partial
// or empty schedule.  sssp_pull_heterogeneous(g, schedule, ...) runs a
// schedule held in these types.
#ifndef HC_CPP_SCHEDULE_H
#define HC_CPP_SCHEDULE_H

#include <algorithm>
#include <ostream>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

#include "../hc_sched.h"
#include "kernels.h"

enum class Device {
    intel_i7_9700K,
    intel_xeon_e5_2686,
    nvidia_quadro_rtx_4000,
    nvidia_tesla_k80,
    nvidia_tesla_m60,
    amd_instinct_mi355x,
    host_cpu,
    undefined
};

constexpr Device DeviceList[] = {Device::intel_i7_9700K,      Device::intel_xeon_e5_2686,
                                 Device::nvidia_quadro_rtx_4000, Device::nvidia_tesla_k80,
                                 Device::nvidia_tesla_m60,    Device::amd_instinct_mi355x,
                                 Device::host_cpu,            Device::undefined};

enum class DeviceType { CPU, GPU, undefined };

namespace hc {
// Device <-> hc_device (same order; undefined has no C id).
inline int32_t device_id_of(Device d) {
    return d == Device::undefined ? -1 : (int32_t)d;
}
}  // namespace hc

inline std::string to_string(Device dev) {
    const int32_t id = hc::device_id_of(dev);
    return id < 0 ? "undefined device" : hc_device_to_string(id);
}

inline Device get_device(const std::string &devstr) {
    for (Device dev : DeviceList)
        if (dev != Device::undefined && devstr == to_string(dev)) return dev;
    return Device::undefined;
}

inline DeviceType get_device_type(Device dev) {
    int32_t t = -1;
    const int32_t id = hc::device_id_of(dev);
    if (id < 0 || hc_device_type_of(id, &t) != HC_OK) return DeviceType::undefined;
    return t == HC_DEV_CPU ? DeviceType::CPU : DeviceType::GPU;
}

inline std::ostream &operator<<(std::ostream &os, Device dev) { return os << to_string(dev); }

struct ScheduleBlock {
    const int device_id;        // instance of its device name in the schedule
    const nid_t start_segment;  // inclusive
    const nid_t end_segment;    // inclusive
    nid_t start_id;             // bookkeeping for a driver
    nid_t end_id;

    ScheduleBlock(int device_id_, nid_t start_segment_, nid_t end_segment_)
        : device_id(device_id_),
          start_segment(start_segment_),
          end_segment(end_segment_),
          start_id(0),
          end_id(0) {}
};

struct SSSPCPUScheduleBlock : public ScheduleBlock {
    const sssp_cpu_epoch_func kernel;
    const SSSPCPU kernel_type;

    SSSPCPUScheduleBlock(int device_id_, nid_t start_segment_, nid_t end_segment_,
                         sssp_cpu_epoch_func kernel_, SSSPCPU kernel_type_)
        : ScheduleBlock(device_id_, start_segment_, end_segment_),
          kernel(kernel_),
          kernel_type(kernel_type_) {}
};

struct SSSPGPUScheduleBlock : public ScheduleBlock {
    const sssp_gpu_epoch_func kernel;
    const SSSPGPU kernel_type;
    const int block_count;
    const int thread_count;

    SSSPGPUScheduleBlock(int device_id_, nid_t start_segment_, nid_t end_segment_,
                         sssp_gpu_epoch_func kernel_, SSSPGPU kernel_type_, int block_count_,
                         int thread_count_)
        : ScheduleBlock(device_id_, start_segment_, end_segment_),
          kernel(kernel_),
          kernel_type(kernel_type_),
          block_count(block_count_),
          thread_count(thread_count_) {}
};

using SSSPScheduleBlock = std::variant<SSSPCPUScheduleBlock, SSSPGPUScheduleBlock>;

struct SSSPHeteroSchedule {
    std::vector<SSSPScheduleBlock> blocks;
    nid_t num_segments = 0;  // segments, not blocks
};

inline std::ostream &operator<<(std::ostream &os, const SSSPCPUScheduleBlock &block) {
    os << "SSSP CPU Block" << std::endl
       << " > device id:     " << block.device_id << std::endl
       << " > start segment: " << block.start_segment << std::endl
       << " > end segment:   " << block.end_segment << std::endl
       << " > kernel:        " << block.kernel_type << std::endl;
    return os;
}

inline std::ostream &operator<<(std::ostream &os, const SSSPGPUScheduleBlock &block) {
    os << "SSSP GPU Block" << std::endl
       << " > device id:     " << block.device_id << std::endl
       << " > start segment: " << block.start_segment << std::endl
       << " > end segment:   " << block.end_segment << std::endl
       << " > kernel:        " << block.kernel_type << std::endl
       << " > block count:   " << block.block_count << std::endl
       << " > thread count:  " << block.thread_count << std::endl;
    return os;
}

inline std::ostream &operator<<(std::ostream &os, const SSSPHeteroSchedule &schedule) {
    os << "SSSP Heterogeneous Schedule" << std::endl
       << " > Number of segments: " << schedule.num_segments << std::endl
       << "Blocks ..." << std::endl;
    for (const SSSPScheduleBlock &block : schedule.blocks)
        std::visit([&os](const auto &b) { os << b; }, block);
    return os;
}

// The registry entry whose name starts `kerstr` (the kernel line of a .skd
// block also carries the launch shape); EnumT::undefined when none does.
template <typename EnumT, typename = std::enable_if_t<std::is_enum<EnumT>::value>>
EnumT get_kernel_id(std::string kerstr) {
    for (EnumT ker : get_kernels(EnumT::undefined)) {
        const std::string name = to_string(ker);
        if (kerstr.compare(0, name.size(), name) == 0) return ker;
    }
    return EnumT::undefined;
}

namespace hc {

inline SSSPCPU sssp_cpu_kernel_of(int32_t id) {
    return id == HC_K_SSSP_CPU_ONE_TO_ONE ? SSSPCPU::one_to_one : SSSPCPU::undefined;
}

inline SSSPGPU sssp_gpu_kernel_of(int32_t id) {
    return id >= HC_K_SSSP_GPU_ONE_TO_ONE && id <= HC_K_SSSP_GPU_BUCKETED
               ? static_cast<SSSPGPU>(id - HC_K_SSSP_GPU_ONE_TO_ONE)
               : SSSPGPU::undefined;
}

// The C schedule of an SSSPHeteroSchedule: CPU blocks on the host cores, GPU
// blocks on this process's GPU (the types keep no device name).
inline hc_schedule_t to_c_schedule(const SSSPHeteroSchedule &schedule) {
    std::vector<hc_schedule_block_t> cb;
    for (const SSSPScheduleBlock &block : schedule.blocks) {
        if (const auto *c = std::get_if<SSSPCPUScheduleBlock>(&block))
            cb.push_back(hc_schedule_block_t{HC_DEVICE_HOST_CPU, c->device_id, c->start_segment,
                                             c->end_segment, kernel_id_of(c->kernel_type), 0, 0});
        else {
            const auto &gb = std::get<SSSPGPUScheduleBlock>(block);
            cb.push_back(hc_schedule_block_t{HC_DEVICE_AMD_MI355X, gb.device_id,
                                             gb.start_segment, gb.end_segment,
                                             kernel_id_of(gb.kernel_type), gb.block_count,
                                             gb.thread_count});
        }
    }
    hc_schedule_t s = nullptr;
    check_or_exit(hc_schedule_create(cb.data(), (int32_t)cb.size(), schedule.num_segments, &s),
                  "hc_schedule_create");
    return s;
}

}  // namespace hc

// Parses a .skd file (a device-name line opens a device instance; then
// "<start> <end>" / "<kernel>[ <blocks> <threads>]" line pairs) into the
// reference's schedule types.
inline SSSPHeteroSchedule load_sssp_schedule(std::string fname) {
    hc_schedule_t s = nullptr;
    hc::check_or_exit(hc_schedule_load(fname.c_str(), &s), "load_sssp_schedule");
    int32_t nb = 0, segs = 0;
    hc::check_or_exit(hc_schedule_info(s, &nb, &segs), "load_sssp_schedule");
    SSSPHeteroSchedule schedule;
    schedule.num_segments = segs;
    for (int32_t i = 0; i < nb; i++) {
        hc_schedule_block_t b{};
        hc::check_or_exit(hc_schedule_get_block(s, i, &b), "load_sssp_schedule");
        int32_t dt = 0;
        hc::check_or_exit(hc_kernel_device_type(b.kernel, &dt), "load_sssp_schedule");
        if (dt == HC_DEV_CPU) {
            const SSSPCPU ker = hc::sssp_cpu_kernel_of(b.kernel);
            if (ker == SSSPCPU::undefined) {
                hc_schedule_free(s);
                std::fprintf(stderr, "load_sssp_schedule: block %d runs \"%s\", not an SSSP kernel\n",
                             i, hc_kernel_to_string(b.kernel));
                std::exit(EXIT_FAILURE);
            }
            schedule.blocks.push_back(SSSPCPUScheduleBlock(b.device_id, b.start_segment,
                                                           b.end_segment, get_kernel(ker), ker));
        } else {
            const SSSPGPU ker = hc::sssp_gpu_kernel_of(b.kernel);
            if (ker == SSSPGPU::undefined) {
                hc_schedule_free(s);
                std::fprintf(stderr, "load_sssp_schedule: block %d runs \"%s\", not an SSSP kernel\n",
                             i, hc_kernel_to_string(b.kernel));
                std::exit(EXIT_FAILURE);
            }
            schedule.blocks.push_back(SSSPGPUScheduleBlock(b.device_id, b.start_segment,
                                                           b.end_segment, get_kernel(ker), ker,
                                                           b.blocks, b.threads));
        }
    }
    hc_schedule_free(s);
    return schedule;
}

// Runs an SSSPHeteroSchedule (sssp_pull_heterogeneous with the schedule in
// memory instead of a file path).
inline double sssp_pull_heterogeneous(const CSRWGraph &g, const SSSPHeteroSchedule &schedule,
                                      const weight_t *init_dist, weight_t **const ret_dist) {
    const bool gpu = std::any_of(schedule.blocks.begin(), schedule.blocks.end(),
                                 [](const SSSPScheduleBlock &b) {
                                     return std::holds_alternative<SSSPGPUScheduleBlock>(b);
                                 });
    if (gpu) hc::check_or_exit(hc_init(0), "hc_init");
    hc_schedule_t s = hc::to_c_schedule(schedule);
    hc::HostView<wnode_t> v(g);
    hc_hetero_t h = nullptr;
    hc::check_or_exit(hc_hetero_create(&v.c, s, &h), "hc_hetero_create");
    weight_t *dist = new weight_t[g.num_nodes];
    hc_run_stats_t st{};
    hc::check_or_exit(hc_sssp_heterogeneous(h, init_dist, dist, &st), "sssp_pull_heterogeneous");
    hc_hetero_free(h);
    hc_schedule_free(s);
    *ret_dist = dist;
    return st.ms;
}

#endif  // HC_CPP_SCHEDULE_H
