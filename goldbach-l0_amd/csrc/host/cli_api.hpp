#pragma once
// The rest of the host API the reference's command-line driver calls
// (src/main.cpp:13-16 includes it from include/l0_context.hpp,
// include/gpu_monitor.hpp and include/sparse_run.hpp), as this build provides
// it: same names, arguments and error behaviour, no Level Zero.
//   sched_policy_name  the EU arbitration policy in force: HIP leaves wave
//                      scheduling to the driver, so "driver"
//                      (include/l0_context.hpp:171).
//   locate_module      the device code is linked into this library: no
//                      module file is looked up, the path comes back as
//                      given (include/l0_context.hpp:109).
//   GpuPower/GpuMonitor the Sysman window (include/gpu_monitor.hpp:22-60):
//                      no power telemetry on this build, status "off" as
//                      the reference reports under GB_SYSMAN=0.
//   run_gpu_sparse     the sparse class map (include/sparse_run.hpp) is not
//                      part of this build: refused with an error.
#include <cstdint>
#include <string>

#include "verifier.hpp"

namespace gb {

const char* sched_policy_name();

std::string locate_module(const std::string& path, bool exe_first = false);

struct GpuPower {
    std::string status = "off";
    double window_s = 0.0;
    bool have_energy = false;
    double energy_j = 0.0;
    double power_w = 0.0;
    bool have_throttle = false;
    double throttled_s = 0.0;
    bool have_freq = false;
    double freq_avg_mhz = 0.0;
    double freq_min_mhz = 0.0;
    uint32_t throttle_reasons = 0;
    uint32_t samples = 0;
    std::string reasons_text() const;
};

class GpuMonitor {
public:
    void start();
    GpuPower stop();

private:
    double t0_ = 0.0;
};

/// @throws std::runtime_error always: the sparse path is not built.
RunStats run_gpu_sparse(uint64_t limit, uint64_t segment_even, const std::string& spv_path, bool verbose,
                        uint64_t start_even, uint32_t sieve_split, const char* prime_cache_path, bool collect_segments);

}  // namespace gb
