// Run-level C ABI (include/goldbach_amd.h) over the C++ driver.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "goldbach_amd.h"
#include "tune.hpp"
#include "verifier.hpp"

namespace {
void put(char* dst, size_t cap, const char* s)
{
    if (!dst || !cap) return;
    std::snprintf(dst, cap, "%s", s);
}
}  // namespace

extern "C" {

const char* gba_version(void) { return gb::VERSION; }

namespace {
void fill(const gb::RunStats& s, gba_run_result* out)
{
        if (out) {
            std::memset(out, 0, sizeof *out);
            out->limit = s.limit;
            out->start = s.start;
            out->segments = s.segments;
            out->leftovers = s.leftovers;
            out->init_seconds = s.init_seconds;
            out->gpu_seconds = s.gpu_seconds;
            out->total_seconds = s.total_seconds;
            out->evens_per_second = s.evens_per_second;
            out->sieve_seconds = s.sieve_seconds;
            out->large_seconds = s.large_seconds;
            out->verify_seconds = s.verify_seconds;
            out->count_seconds = s.count_seconds;
            out->sync_seconds = s.sync_seconds;
            out->total_tiles = s.total_tiles;
            out->total_large_items = s.total_large_items;
            out->split_used = s.split_used;
            out->stats_on = s.stats_on;
            out->stats_pmax = s.stats_pmax;
            out->stats_primes = s.stats_primes;
            out->stats_twins = s.stats_twins;
            out->stats_sampled = s.stats_sampled;
            out->stats_tail = s.stats_tail;
            out->harvest_kernel_seconds = s.harvest_kernel_seconds;
            out->window_kernel_seconds = s.window_kernel_seconds;
            out->harvest_launches = s.harvest_launches;
            out->window_launches = s.window_launches;
            out->harvest_bytes = s.harvest_bytes;
            out->window_bytes = s.window_bytes;
            out->host_plan_seconds = s.host_plan_seconds;
            out->harvest_reruns = s.harvest_reruns;
            out->graph_launches = s.graph_launches;
            out->harvest_words = s.harvest_words;
        }
}
}  // namespace

struct gba_engine {
    gb::Engine e;
    gba_engine(uint64_t l, uint64_t s) : e(l, s) {}
};

int gba_engine_open(uint64_t max_limit, uint64_t max_span, gba_engine** out, double* init_seconds, char* err,
                    size_t err_cap)
{
    try {
        *out = new gba_engine(max_limit, max_span);
        if (init_seconds) *init_seconds = (*out)->e.init_seconds();
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        *out = nullptr;
        put(err, err_cap, ex.what());
        return 1;
    }
}

int gba_engine_run(gba_engine* e, uint64_t limit, uint64_t segment_even, uint64_t start_even, gba_run_result* out,
                   char* stats_line, size_t stats_cap, char* err, size_t err_cap)
{
    try {
        const gb::RunStats s = e->e.run(limit, segment_even, false, start_even);
        fill(s, out);
        put(stats_line, stats_cap, s.stats_line.c_str());
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        put(err, err_cap, ex.what());
        return 1;
    }
}

void gba_engine_close(gba_engine* e) { delete e; }

int gba_engine_prepare(gba_engine* e, uint64_t limit, uint64_t start_even, double* seconds, char* err, size_t err_cap)
{
    try {
        const double s = e->e.prepare(limit, start_even);
        if (seconds) *seconds = s;
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        put(err, err_cap, ex.what());
        return 1;
    }
}

int gba_run_gpu(uint64_t limit, uint64_t segment_even, const char* spv_path, int verbose, uint64_t start_even,
                uint32_t sieve_split, const char* prime_cache_path, int collect_segments, gba_run_result* out,
                gba_segment_stats* segs, size_t seg_cap, size_t* n_segs, char* stats_line, size_t stats_cap,
                char* err, size_t err_cap)
{
    try {
        const gb::RunStats s = gb::run_gpu(limit, segment_even, spv_path ? spv_path : "", verbose != 0, start_even,
                                           sieve_split, prime_cache_path, collect_segments != 0);
        fill(s, out);
        if (n_segs) *n_segs = s.per_segment.size();
        for (size_t k = 0; segs && k < std::min(seg_cap, s.per_segment.size()); ++k) {
            const gb::SegmentStats& g = s.per_segment[k];
            gba_segment_stats& d = segs[k];
            std::memset(&d, 0, sizeof d);
            d.seg_start = g.seg_start;
            d.seg_end = g.seg_end;
            d.evens = g.evens;
            d.tiles = g.tiles;
            d.large_items = g.large_items;
            d.sieve_split = g.sieve_split;
            d.mask_count = g.mask_count;
            d.leftovers = g.leftovers;
            d.sieve_seconds = g.sieve_seconds;
            d.large_seconds = g.large_seconds;
            d.check_seconds = g.check_seconds;
            d.offset_seconds = g.offset_seconds;
            d.wall_seconds = g.wall_seconds;
        }
        put(stats_line, stats_cap, s.stats_line.c_str());
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& e) {
        put(err, err_cap, e.what());
        return 1;
    }
}

int gba_segment_map(uint64_t seg_start, uint64_t limit, uint64_t span, uint64_t* words, size_t cap,
                    uint32_t* cw_out, uint64_t* Q, uint32_t* nc, char* err, size_t err_cap)
{
    try {
        std::vector<uint64_t> w;
        uint64_t q[8];
        uint32_t n[8];
        const uint32_t cw = gb::gpu_segment_map(seg_start, limit, span, w, q, n);
        if (cw_out) *cw_out = cw;
        if (words) std::memcpy(words, w.data(), std::min(cap, w.size()) * 8);
        for (int k = 0; k < 8; ++k) {
            if (Q) Q[k] = q[k];
            if (nc) nc[k] = n[k];
        }
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& e) {
        put(err, err_cap, e.what());
        return 1;
    }
}

struct gba_batcher {
    gb::Batcher b;
    std::vector<size_t> finished;   // completed ranges not yet reported
    gba_batcher(gba_engine* e, uint64_t max_limit, uint64_t min_start, uint64_t span, uint64_t budget)
        : b(e->e, max_limit, min_start, span, budget)
    {
    }
};

int gba_batcher_open(gba_engine* e, uint64_t max_limit, uint64_t min_start, uint64_t segment_even,
                     uint64_t device_budget, gba_batcher** out, uint32_t* batch_segments, char* err, size_t err_cap)
{
    try {
        *out = new gba_batcher(e, max_limit, min_start, segment_even, device_budget);
        if (batch_segments) *batch_segments = (*out)->b.batch_segments();
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        *out = nullptr;
        put(err, err_cap, ex.what());
        return 1;
    }
}

int gba_batcher_submit(gba_batcher* b, uint64_t start_even, uint64_t limit, uint64_t* id, char* err, size_t err_cap)
{
    try {
        const size_t i = b->b.submit(start_even, limit);
        if (id) *id = i;
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        put(err, err_cap, ex.what());
        return 1;
    }
}

int gba_batcher_step(gba_batcher* b, uint64_t* finished, uint32_t finished_cap, uint32_t* n_finished, int* more,
                     char* err, size_t err_cap)
{
    try {
        const bool busy = b->b.step(&b->finished);
        const uint32_t n = uint32_t(std::min<size_t>(finished ? finished_cap : 0, b->finished.size()));
        for (uint32_t k = 0; k < n; ++k) finished[k] = b->finished[k];
        b->finished.erase(b->finished.begin(), b->finished.begin() + n);
        if (n_finished) *n_finished = n;
        if (more) *more = busy || !b->finished.empty();
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        put(err, err_cap, ex.what());
        return 1;
    }
}

int gba_batcher_result(gba_batcher* b, uint64_t id, gba_run_result* out, char* stats_line, size_t stats_cap,
                       char* state, size_t state_cap, char* err, size_t err_cap)
{
    try {
        const gb::Batcher::Result& r = b->b.result(size_t(id));
        if (!r.done) throw std::runtime_error("range " + std::to_string(id) + " not finished");
        fill(r.stats, out);
        put(stats_line, stats_cap, r.stats.stats_line.c_str());
        put(state, state_cap, r.state.c_str());
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        put(err, err_cap, ex.what());
        return 1;
    }
}

int gba_batcher_totals(gba_batcher* b, gba_run_result* out)
{
    fill(b->b.totals(), out);
    return 0;
}

int gba_batcher_memory(gba_batcher* b, uint64_t* device_bytes, uint32_t* inflight_segments)
{
    if (device_bytes) *device_bytes = b->b.device_bytes();
    if (inflight_segments) *inflight_segments = 2 * b->b.batch_segments();
    return 0;
}

int gba_batcher_segments(gba_batcher* b, uint64_t start_even, uint64_t limit, uint64_t* segments, char* err,
                         size_t err_cap)
{
    try {
        *segments = b->b.segments_of(start_even, limit);
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        put(err, err_cap, ex.what());
        return 1;
    }
}

void gba_batcher_close(gba_batcher* b) { delete b; }

int gba_tune_load(const char* path, uint64_t limit, char* applied, size_t applied_cap, char* err, size_t err_cap)
{
    try {
        const std::string a = path ? gb::tune_load(path, limit) : (gb::tune_clear(), std::string());
        put(applied, applied_cap, a.c_str());
        put(err, err_cap, "");
        return 0;
    } catch (const std::exception& ex) {
        put(err, err_cap, ex.what());
        return 1;
    }
}

}  // extern "C"
