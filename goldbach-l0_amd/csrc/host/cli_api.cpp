// The reference CLI's remaining host API on this build (cli_api.hpp).
#include "cli_api.hpp"

#include <chrono>
#include <stdexcept>

namespace gb {

const char* sched_policy_name() { return "driver"; }

std::string locate_module(const std::string& path, bool) { return path; }

std::string GpuPower::reasons_text() const { return "none"; }

namespace {
double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void GpuMonitor::start() { t0_ = now_s(); }

GpuPower GpuMonitor::stop()
{
    GpuPower p;
    p.window_s = now_s() - t0_;
    return p;
}

RunStats run_gpu_sparse(uint64_t, uint64_t, const std::string&, bool, uint64_t, uint32_t, const char*, bool)
{
    throw std::runtime_error("--sparse (the sparse class-map path) is not part of this build");
}

}  // namespace gb
