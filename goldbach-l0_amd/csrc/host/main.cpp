// goldbach_amd — command-line verifier (the reference's src/main.cpp surface:
// the same options, number syntax, exit codes 0 / 1 / 2 and output formats,
// with the device block reporting this build's GPU).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <optional>
#include <string>
#include <string_view>

#include "config.hpp"
#include "env.hpp"
#include "textio.hpp"
#include "cli_api.hpp"
#include "verifier.hpp"

namespace {

/// A count on the command line: full decimal or e-notation ("4e18"), the
/// whole string, exact; nullopt for anything else or a value above 2^64 - 1
/// (the reference's number contract, src/main.cpp:28-57 of the reference:
/// a typo is a usage error, never a silently truncated limit).
std::optional<uint64_t> parse_count(std::string_view s)
{
    const size_t e = s.find_first_of("eE");
    const std::string_view mantissa = s.substr(0, e);
    gb::textio::u128 v = 0;
    if (!gb::textio::parse(mantissa, v)) return std::nullopt;
    if (e != std::string_view::npos) {
        uint32_t exponent = 0;
        if (!gb::textio::parse(s.substr(e + 1), exponent) || exponent > 20) return std::nullopt;
        // stop once above 2^64 - 1: each product then stays far below 2^128
        while (exponent-- && v <= UINT64_MAX) v *= 10;
    }
    if (v > UINT64_MAX) return std::nullopt;
    return uint64_t(v);
}

void usage()
{
    std::printf(
        "goldbach-l0 (MI355X) — Goldbach verifier for AMD Instinct MI355X (HIP, gfx950)\n"
        "\n"
        "usage: goldbach_amd [--limit N] [--start N] [--segment N] [--split N]\n"
        "                    [--spv FILE] [--prime-cache FILE]\n"
        "                    [--cpu-check N] [--json] [--csv] [--per-seg]\n"
        "                    [--info] [--verbose]\n"
        "\n"
        "  N accepts full decimal or e-notation: 4000000000000000000 = 4e18.\n"
        "\n"
        "  --limit N      verify every even n in [start, N]  (default %llu)\n"
        "  --start N      first even n to verify (scale tests; default 4)\n"
        "  --segment N    even integer span per GPU segment  (default %llu)\n"
        "  --split N      base-prime split index (0 = auto)  (default 0)\n"
        "  --spv FILE     accepted for compatibility (the device code is linked in)\n"
        "  --prime-cache FILE  GBP2 base-prime cache stem: FILE.<bound> (default base_primes.cache)\n"
        "  --cpu-check N  run the independent CPU check up to N and exit\n"
        "  --json         machine-readable JSON report on stdout\n"
        "  --csv          one summary line: key=value pairs (grep/awk friendly)\n"
        "  --per-seg      print a per-segment timing table (implies collection)\n"
        "  --info         print device/driver/config info and exit\n"
        "  --verbose      per-segment progress output\n",
        (unsigned long long)gb::DEFAULT_LIMIT, (unsigned long long)gb::DEFAULT_SEGMENT_EVEN);
}


void print_csv(const gb::RunStats& s, uint64_t segment, uint32_t split)
{
    std::printf(
        "version=%s limit=%llu start=%llu segment=%llu split=%u segments=%llu leftovers=%llu result=PASS "
        "init_s=%.4f init_cache_s=%.4f init_module_s=%.4f init_upload_s=%.4f gpu_s=%.4f total_s=%.4f sieve_s=%.4f "
        "large_s=%.4f check_s=%.4f offsets_s=%.4f table_s=%.4f sync_s=%.4f evens_per_sec=%.6g tiles=%llu "
        "large_items=%llu ridge_min=%u ridge_found=%llu ridge_dropped=%llu ridge_max=%llu stats=%d stats_tail=%llu "
        "stats_sampled=%llu stats_pmax=%llu primes=%llu twins=%llu split_used=%u sched=%s sysman=off\n",
        gb::VERSION, (unsigned long long)s.limit, (unsigned long long)s.start, (unsigned long long)segment, split,
        (unsigned long long)s.segments, (unsigned long long)s.leftovers, s.init_seconds, s.init_cache_s,
        s.init_module_s, s.init_upload_s, s.gpu_seconds, s.total_seconds, s.sieve_seconds, s.large_seconds,
        s.verify_seconds, s.offset_seconds, s.table_seconds, s.sync_seconds, s.evens_per_second,
        (unsigned long long)s.total_tiles, (unsigned long long)s.total_large_items, s.ridge_min,
        (unsigned long long)s.ridge_found, (unsigned long long)s.ridge_dropped,
        (unsigned long long)(s.ridge_top.empty() ? 0 : s.ridge_top.front().p), s.stats_on ? 1 : 0,
        (unsigned long long)s.stats_tail, (unsigned long long)s.stats_sampled, (unsigned long long)s.stats_pmax,
        (unsigned long long)s.stats_primes, (unsigned long long)s.stats_twins, s.split_used, gb::sched_policy_name());
}

void print_json(const gb::RunStats& s, uint64_t segment, uint32_t split)
{
    std::printf("{\n  \"version\": \"%s\",\n  \"limit\": %llu,\n  \"start\": %llu,\n  \"segment\": %llu,\n"
                "  \"split\": %u,\n", gb::VERSION, (unsigned long long)s.limit, (unsigned long long)s.start,
                (unsigned long long)segment, split);
    if (s.split_used) std::printf("  \"split_used\": %u,\n", s.split_used);
    std::printf("  \"sched\": \"%s\",\n  \"spv\": \"\",\n  \"segments\": %llu,\n  \"leftovers\": %llu,\n"
                "  \"result\": \"PASS\",\n", gb::sched_policy_name(), (unsigned long long)s.segments,
                (unsigned long long)s.leftovers);
    std::printf("  \"init_seconds\": %.4f,\n  \"gpu_seconds\": %.4f,\n  \"total_seconds\": %.4f,\n"
                "  \"sieve_seconds\": %.4f,\n  \"large_seconds\": %.4f,\n  \"verify_seconds\": %.4f,\n"
                "  \"offset_seconds\": %.4f,\n  \"table_seconds\": %.4f,\n  \"sync_seconds\": %.4f,\n"
                "  \"evens_per_second\": %.6g,\n  \"total_tiles\": %llu,\n  \"total_large_items\": %llu,\n"
                "  \"ridge_min\": %u,\n  \"ridge_found\": %llu,\n  \"ridge_dropped\": %llu,\n  \"ridge_top\": [",
                s.init_seconds, s.gpu_seconds, s.total_seconds, s.sieve_seconds, s.large_seconds, s.verify_seconds,
                s.offset_seconds, s.table_seconds, s.sync_seconds, s.evens_per_second,
                (unsigned long long)s.total_tiles, (unsigned long long)s.total_large_items, s.ridge_min,
                (unsigned long long)s.ridge_found, (unsigned long long)s.ridge_dropped);
    for (size_t i = 0; i < s.ridge_top.size(); ++i) {
        const auto& e = s.ridge_top[i];
        std::printf("%s\n    {\"rank\": %zu, \"ridge\": %llu, \"even\": %llu, \"p\": %llu, \"q\": %llu, "
                    "\"segment\": %llu}", i ? "," : "", i + 1, (unsigned long long)e.p, (unsigned long long)e.n,
                    (unsigned long long)e.p, (unsigned long long)e.q, (unsigned long long)e.seg);
    }
    std::printf("%s]", s.ridge_top.empty() ? "" : "\n  ");
    std::printf(",\n  \"sysman\": {\"status\": \"off\", \"window_s\": 0.0000}");
    if (s.stats_on)
        std::printf(",\n  \"stats\": {\"tail_evens\": %llu, \"sampled_evens\": %llu, \"pmax\": %llu, \"primes\": %llu, "
                    "\"twins\": %llu}", (unsigned long long)s.stats_tail, (unsigned long long)s.stats_sampled,
                    (unsigned long long)s.stats_pmax, (unsigned long long)s.stats_primes,
                    (unsigned long long)s.stats_twins);
    if (!s.per_segment.empty()) {
        std::printf(",\n  \"per_segment\": [\n");
        for (size_t i = 0; i < s.per_segment.size(); ++i) {
            const auto& g = s.per_segment[i];
            std::printf("    {\"seg\": %zu, \"start\": %llu, \"end\": %llu, \"evens\": %llu, \"tiles\": %llu, "
                        "\"large_items\": %llu, \"split\": %u, \"mask_count\": %u, \"leftovers\": %u, "
                        "\"sieve_s\": %.4f, \"large_s\": %.4f, \"check_s\": %.4f, \"offset_s\": %.4f, "
                        "\"wall_s\": %.4f}%s\n", i + 1, (unsigned long long)g.seg_start,
                        (unsigned long long)g.seg_end, (unsigned long long)g.evens, (unsigned long long)g.tiles,
                        (unsigned long long)g.large_items, g.sieve_split, g.mask_count, g.leftovers, g.sieve_seconds,
                        g.large_seconds, g.check_seconds, g.offset_seconds, g.wall_seconds,
                        i + 1 < s.per_segment.size() ? "," : "");
        }
        std::printf("  ]");
    }
    std::printf("\n}\n");
}

// The persistent ridge leaderboard (the reference's print_ridges,
// src/main.cpp:244-256): the largest recorded ridges, descending.
void print_ridges(const gb::RunStats& s)
{
    if (!s.ridge_min) return;
    std::printf("\nMy Top 30 Largest Goldbach Ridges\n");
    if (s.ridge_top.empty()) {
        std::printf("  (none recorded at or above the floor %u yet)\n", s.ridge_min);
        return;
    }
    for (size_t i = 0; i < s.ridge_top.size(); ++i) {
        const auto& e = s.ridge_top[i];
        std::printf("  %2zu. ridge %-5llu  %llu = %llu + %llu\n", i + 1, (unsigned long long)e.p,
                    (unsigned long long)e.n, (unsigned long long)e.q, (unsigned long long)e.p);
    }
}

void print_per_segment(const gb::RunStats& s)
{
    std::printf("#seg  evens  [start, end]      tiles  large_items  split  sieve_s  large_s  check_s  offset_s  wall_s  left\n");
    for (size_t i = 0; i < s.per_segment.size(); ++i) {
        const auto& g = s.per_segment[i];
        std::printf("%4zu  %9llu  [%13llu, %-13llu] %6llu %11llu  %6u  %6.3f  %7.3f  %7.3f  %7.3f  %6.3f  %5u\n", i + 1,
                    (unsigned long long)g.evens, (unsigned long long)g.seg_start, (unsigned long long)g.seg_end,
                    (unsigned long long)g.tiles, (unsigned long long)g.large_items, g.sieve_split, g.sieve_seconds,
                    g.large_seconds, g.check_seconds, g.offset_seconds, g.wall_seconds, g.leftovers);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    uint64_t limit = gb::DEFAULT_LIMIT, start = 4, segment = gb::DEFAULT_SEGMENT_EVEN, split = 0, cpu_check = 0;
    bool cpu_check_set = false, verbose = false, json = false, csv = false, per_seg = false;
    std::string spv, prime_cache = "base_primes.cache";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const auto next = [&](uint64_t& out) {
            if (i + 1 >= argc) { usage(); std::exit(2); }
            const std::optional<uint64_t> v = parse_count(argv[++i]);
            if (!v) {
                std::fprintf(stderr, "error: invalid number '%s'\n", argv[i]);
                std::exit(2);
            }
            out = *v;
        };
        const auto next_str = [&]() -> const char* {
            if (i + 1 >= argc) { usage(); std::exit(2); }
            return argv[++i];
        };
        if (a == "--limit") next(limit);
        else if (a == "--start") next(start);
        else if (a == "--segment") next(segment);
        else if (a == "--split") next(split);
        else if (a == "--cpu-check") { next(cpu_check); cpu_check_set = true; }
        else if (a == "--spv") spv = next_str();
        else if (a == "--prime-cache") prime_cache = next_str();
        else if (a == "--verbose") verbose = true;
        else if (a == "--json") json = true;
        else if (a == "--csv") csv = true;
        else if (a == "--per-seg") per_seg = true;
        else if (a == "--sparse") {
            std::fprintf(stderr, "error: --sparse (the sparse class-map path) is not part of this build\n");
            return 2;
        } else if (a == "--info") {
            try {
                gb::print_device_info();
            } catch (const std::exception& e) {
                std::fprintf(stderr, "error: %s\n", e.what());
                return 1;
            }
            return 0;
        } else if (a == "--help" || a == "-h") { usage(); return 0; }
        else { std::fprintf(stderr, "unknown option: %s\n", a.c_str()); usage(); return 2; }
    }
    if (cpu_check_set && cpu_check < 4) {
        std::fprintf(stderr, "error: --cpu-check needs N >= 4\n");
        return 2;
    }
    if (!cpu_check_set) {
        const uint64_t first_even = start < 4 ? 4 : start + (start & 1);
        if (start > limit || first_even > limit || first_even < start) {
            std::fprintf(stderr, "error: no even n >= 4 in [start=%llu, limit=%llu]\n", (unsigned long long)start,
                         (unsigned long long)limit);
            return 2;
        }
        if (split > 0xFFFFFFFFull) {
            std::fprintf(stderr, "error: --split %llu exceeds 2^32 - 1\n", (unsigned long long)split);
            return 2;
        }
    }
    try {
        if (cpu_check_set) {
            const bool ok = gb::verify_cpu(cpu_check, verbose);
            std::printf("CPU check up to %llu: %s\n", (unsigned long long)cpu_check, ok ? "PASS" : "FAIL");
            return ok ? 0 : 1;
        }
        if (gb::env::onoff("GB_SPARSE", false)) {
            std::fprintf(stderr, "error: GB_SPARSE (the sparse class-map path) is not part of this build\n");
            return 2;
        }
        const gb::RunStats s = gb::run_gpu(limit, segment, spv, verbose, start, uint32_t(split), prime_cache.c_str(),
                                           per_seg || json);
        if (json) print_json(s, segment, uint32_t(split));
        else if (csv) print_csv(s, segment, uint32_t(split));
        else
            std::printf("goldbach-l0 v%s\n  limit          : %llu\n  start          : %llu\n  segments       : %llu\n"
                        "  leftovers      : %llu\n  init time      : %.3f s\n  gpu time       : %.3f s\n"
                        "    sieve        : %.3f s\n    large sieve  : %.3f s\n    verify       : %.3f s\n"
                        "    offsets      : %.3f s\n    sync/other   : %.3f s\n  total          : %.3f s\n"
                        "  evens/second   : %.3g\n  EU arbitration : %s\n  GPU (sysman)   : off\n"
                        "  result         : PASS\n",
                        gb::VERSION, (unsigned long long)s.limit, (unsigned long long)s.start,
                        (unsigned long long)s.segments, (unsigned long long)s.leftovers, s.init_seconds,
                        s.gpu_seconds, s.sieve_seconds, s.large_seconds, s.verify_seconds, s.offset_seconds,
                        s.sync_seconds, s.total_seconds, s.evens_per_second, gb::sched_policy_name());
        if (s.stats_on && !csv && !json)
            std::printf("\nGoldbach statistics (exact; log %s, state %s)\n  largest p(n)            : %llu\n"
                        "  evens with p(n) >= tail : %llu\n  evens with full p(n)    : %llu (sampled segments)\n"
                        "  primes in [start, limit]: %llu\n  twin-prime pairs        : %llu\n",
                        gb::env::path("GB_STATS_LOG", "goldbach_stats.log").c_str(),
                        gb::env::path("GB_STATS_STATE", "goldbach_stats.state").c_str(),
                        (unsigned long long)s.stats_pmax, (unsigned long long)s.stats_tail,
                        (unsigned long long)s.stats_sampled, (unsigned long long)s.stats_primes,
                        (unsigned long long)s.stats_twins);
        if (!csv && !json) print_ridges(s);   // --json / --csv carry ridge_top / ridge_max
        if (per_seg && !json) print_per_segment(s);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
