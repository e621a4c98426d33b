#pragma once
// GPU verification driver of the dense path: the reference's host-level API
// (include/verifier.hpp:12-123: SegmentStats, RunStats, run_gpu,
// print_device_info, verify_cpu), same names, arguments and error behaviour,
// running the MI355X pipeline through the gbh C ABI (include/gbh.h).
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace gb {

struct SegmentStats {
    uint64_t seg_start = 0, seg_end = 0, evens = 0;
    uint64_t tiles = 0;          // window work-groups of the window sieve
    uint64_t large_items = 0;    // large-prime work-items
    uint32_t sieve_split = 0;
    uint32_t mask_count = 0;
    uint32_t leftovers = 0;
    double sieve_seconds = 0, large_seconds = 0, check_seconds = 0, offset_seconds = 0, wall_seconds = 0;
};

struct RidgeEntry {
    uint64_t n = 0, p = 0, q = 0, seg = 0, found = 0;
};

struct RunStats {
    uint64_t limit = 0, start = 0, segment_used = 0, segments = 0, leftovers = 0;
    double init_seconds = 0, init_cache_s = 0, init_module_s = 0, init_upload_s = 0;
    double gpu_seconds = 0, total_seconds = 0, evens_per_second = 0;
    double sieve_seconds = 0, large_seconds = 0, verify_seconds = 0, sync_seconds = 0;
    double offset_seconds = 0, table_seconds = 0;
    double count_seconds = 0;    // counter + gap walk (part of verify_seconds)
    uint64_t total_tiles = 0, total_large_items = 0;
    uint32_t split_used = 0;
    uint32_t ridge_min = 0;
    uint64_t ridge_found = 0, ridge_dropped = 0;
    std::vector<RidgeEntry> ridge_top;
    bool stats_on = false;
    uint64_t stats_pmax = 0, stats_primes = 0, stats_twins = 0, stats_sampled = 0, stats_tail = 0;
    std::string stats_line;      // the run's GBSTATS line (empty: stats off)
    // device time of the hit-list harvest and window sieve kernels alone
    // (HIP events on the run's stream around each launch), their launches
    // and their algorithmic HBM bytes over the run (the data each must move
    // at least once: primes, bitmap words and lists written for the
    // harvest; tables and lists read, map written for the window sieve)
    double harvest_kernel_seconds = 0, window_kernel_seconds = 0;
    uint64_t harvest_launches = 0, window_launches = 0;
    double harvest_bytes = 0, window_bytes = 0;
    double host_plan_seconds = 0;   // range resolution, statistics set-up and buffer planning of this run
    uint64_t harvest_reruns = 0;    // batches re-run with counted lists after a direct-harvest region overflow
    uint64_t graph_launches = 0;    // batches launched as one captured HIP graph (GB_GRAPH)
    double harvest_words = 0;       // of harvest_bytes: the 8-byte multiplier-bitmap words read (modelled)
    std::vector<SegmentStats> per_segment;
};

/// Verifies Goldbach for every even n in [max(start_even, 4) rounded up to
/// even, limit] on the GPU (verifier.hpp:71-110 of the reference).
/// @param spv_path  accepted for interface compatibility; the device code is
///                  linked into this library (the argument is ignored).
/// @throws std::runtime_error on any HIP error, device error flag, overflow,
///         coverage mismatch, statistics file problem or a Goldbach
///         counterexample; nothing is reported as PASS then.
RunStats run_gpu(uint64_t limit, uint64_t segment_even, const std::string& spv_path, bool verbose,
                 uint64_t start_even = 4, uint32_t sieve_split_arg = 0, const char* prime_cache_path = nullptr,
                 bool collect_segments = false);

/// A verifier with its device state resident between runs: the base primes
/// up to sqrt(max_limit), the prime tables on the GPU and the segment
/// buffers for spans up to max_span. run() then verifies any range below
/// max_limit without repeating that set-up (serving and batched callers;
/// run_gpu is one Engine, one run).
class Engine {
public:
    /// @throws std::runtime_error when no gfx950 device is usable.
    /// @param prime_cache GBP2 base-prime cache path stem (nullptr: none)
    Engine(uint64_t max_limit, uint64_t max_span = 467'500'000ULL, const char* prime_cache = nullptr);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    /// Same contract as run_gpu; limit <= max_limit and the resolved
    /// segment span <= max_span.
    RunStats run(uint64_t limit, uint64_t segment_even, bool verbose = false, uint64_t start_even = 4,
                 uint32_t sieve_split_arg = 0, bool collect_segments = false);
    /// Builds the range-dependent resident tables (the multiplier bitmap)
    /// for runs over [start_even, limit]; returns the seconds the build took
    /// (0 when they were resident).
    double prepare(uint64_t limit, uint64_t start_even = 4);
    double init_seconds() const;
    /// init breakdown: base primes, device context, uploads and buffers
    void init_times(double& cache, double& module, double& upload) const;
    uint64_t max_limit() const;
    struct State;
    State* state() { return s_.get(); }

private:
    std::unique_ptr<State> s_;
};

/// Many ranges through one engine's pipeline (batched and served
/// verification; no counterpart in the reference, which verifies one range
/// per process). Each batch packs the waiting segments of several ranges
/// (each range split into segments exactly as a plain run splits it), one
/// launch per stage for the whole batch; the large-prime harvest walks each
/// range's q range separately. Every range's statistics, ridges and
/// fail-closed checks are those of a plain run of that range alone, without
/// campaign files: the result holds the GBSTATS line and the state text a
/// fresh campaign would hold after the run. One thread per Batcher.
class Batcher {
public:
    struct Result {
        RunStats stats;          // stats_line: the GBSTATS line
        std::string state;       // the state file text of a fresh campaign after this range alone
        bool done = false;
    };
    /// Plans the pipeline (large-prime split, multiplier bitmap, buffers)
    /// for ranges with start >= @p min_start and limit <= @p max_limit. The
    /// batcher's batch buffers are its own, allocated here: with a
    /// @p device_budget (bytes, 0: none) the segments per batch are the most
    /// whose buffers fit in 7/8 of it, and the direct harvest's regions are
    /// allocated only where they fit in the rest (else the counted form), so
    /// device_bytes() never exceeds the budget.
    /// @throws std::runtime_error when the bounds exceed the engine's, or
    /// one-segment batches do not fit the budget.
    Batcher(Engine& e, uint64_t max_limit, uint64_t min_start, uint64_t segment_even = 467'500'000ULL,
            uint64_t device_budget = 0);
    ~Batcher();
    Batcher(const Batcher&) = delete;
    Batcher& operator=(const Batcher&) = delete;
    /// Queues [start_even, limit]; @return its id (0, 1, ...).
    /// @throws std::runtime_error for a range outside the planned bounds.
    size_t submit(uint64_t start_even, uint64_t limit);
    /// Issues the next batch of waiting segments (if any) and collects the
    /// batch issued before it; appends the ids of ranges completed by that
    /// collection to @p finished. @return false when nothing waits or is in
    /// flight. @throws std::runtime_error on any device error or failed
    /// check (the Batcher is unusable afterwards).
    bool step(std::vector<size_t>* finished = nullptr);
    const Result& result(size_t id) const;
    size_t size() const;
    /// Segments a batch may hold (two batches are in flight at once).
    uint32_t batch_segments() const;
    /// Device bytes this batcher's batch buffers hold now.
    uint64_t device_bytes() const;
    /// Segments [start_even, limit] splits into (as a plain run splits it).
    /// @throws std::runtime_error for a range outside the engine's bounds.
    uint64_t segments_of(uint64_t start_even, uint64_t limit) const;
    /// Kernel times and algorithmic bytes over every batch so far.
    const RunStats& totals() const;

private:
    struct Impl;
    std::unique_ptr<Impl> p_;
};

void print_device_info();

/// Independent CPU check by trial division for every even n in [4, limit].
bool verify_cpu(uint64_t limit, bool verbose);

/// One segment's class map built on the GPU, copied back (8 class arrays of
/// @p words_out words each, stride @p cw words); a validator entry point
/// (the reference's goldbach-tools debug_map role).
/// @return the class stride cw; fills Q[8] and nc[8].
uint32_t gpu_segment_map(uint64_t seg_start, uint64_t limit, uint64_t span, std::vector<uint64_t>& words,
                         uint64_t Q[8], uint32_t nc[8]);

}  // namespace gb
