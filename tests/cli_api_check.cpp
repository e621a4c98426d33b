// Compile-and-run check of the host API the reference's command-line driver
// (src/main.cpp of the reference) uses, against this build's headers and
// library: every gb:: name, RunStats / per-segment / ridge field and
// GpuPower field that driver reads exists with a compatible type, and the
// calls that need no GPU behave as documented (cli_api.hpp). Built and run
// by tests/test_host.py.
#include <cstdio>
#include <stdexcept>
#include <string>

#include "cli_api.hpp"
#include "config.hpp"
#include "env.hpp"
#include "verifier.hpp"

namespace {

int failed = 0;

void expect(bool ok, const char* what)
{
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++failed;
    }
}

// every field the driver prints, read once (types must convert as it uses them)
double touch(const gb::RunStats& s)
{
    double t = double(s.limit + s.start + s.segments + s.leftovers + s.total_tiles + s.total_large_items +
                      s.ridge_min + s.ridge_found + s.ridge_dropped + s.segment_used + s.split_used + s.stats_tail +
                      s.stats_sampled + s.stats_pmax + s.stats_primes + s.stats_twins + (s.stats_on ? 1 : 0));
    t += s.init_seconds + s.init_cache_s + s.init_module_s + s.init_upload_s + s.gpu_seconds + s.total_seconds +
         s.sieve_seconds + s.large_seconds + s.verify_seconds + s.offset_seconds + s.table_seconds + s.sync_seconds +
         s.evens_per_second;
    for (const auto& e : s.ridge_top) t += double(e.p + e.n + e.q + e.seg);
    for (const auto& g : s.per_segment)
        t += double(g.seg_start + g.seg_end + g.evens + g.tiles + g.large_items + g.sieve_split + g.mask_count +
                    g.leftovers) +
             g.sieve_seconds + g.large_seconds + g.check_seconds + g.offset_seconds + g.wall_seconds;
    return t;
}

}  // namespace

int main()
{
    expect(std::string(gb::sched_policy_name()) == "driver", "sched_policy_name");
    expect(gb::locate_module("kernels.spv", true) == "kernels.spv", "locate_module");
    gb::GpuMonitor monitor;
    monitor.start();
    const gb::GpuPower pw = monitor.stop();
    expect(pw.status == "off" && !pw.have_energy && !pw.have_freq && !pw.have_throttle && pw.window_s >= 0,
           "GpuMonitor");
    expect(pw.reasons_text() == "none" && pw.samples == 0 && pw.energy_j == 0 && pw.power_w == 0 &&
               pw.throttled_s == 0 && pw.freq_avg_mhz == 0 && pw.freq_min_mhz == 0,
           "GpuPower fields");
    bool refused = false;
    try {
        (void)gb::run_gpu_sparse(gb::DEFAULT_LIMIT, gb::DEFAULT_SEGMENT_EVEN, "", false, 4, 0, "", false);
    } catch (const std::runtime_error& e) {
        refused = std::string(e.what()).find("sparse") != std::string::npos;
    }
    expect(refused, "run_gpu_sparse refuses");
    expect(gb::verify_cpu(100000, false), "verify_cpu");
    expect(gb::env::onoff("GB_CLI_API_CHECK_UNSET", true) && !gb::env::present("GB_CLI_API_CHECK_UNSET") &&
               gb::env::raw("GB_CLI_API_CHECK_UNSET") == nullptr,
           "env");
    expect(std::string(gb::VERSION).size() > 0, "VERSION");
    gb::RunStats s;
    expect(touch(s) == 0, "RunStats fields");
    // the GPU entry points the driver calls (addresses only: no GPU here)
    gb::RunStats (*run)(uint64_t, uint64_t, const std::string&, bool, uint64_t, uint32_t, const char*, bool) =
        &gb::run_gpu;
    expect(run != nullptr, "run_gpu");
    void (*info)() = &gb::print_device_info;
    expect(info != nullptr, "print_device_info");
    std::printf("cli api check: %d failed\n", failed);
    return failed ? 1 : 0;
}
