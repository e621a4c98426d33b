// Run-level C ABI (include/goldbach_amd.h) over the C++ driver.
#include <cstdio>
#include <cstring>
#include <exception>

#include "config.hpp"
#include "goldbach_amd.h"
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

int gba_run_gpu(uint64_t limit, uint64_t segment_even, int verbose, uint64_t start_even, uint32_t sieve_split,
                gba_run_result* out, char* stats_line, size_t stats_cap, char* err, size_t err_cap)
{
    try {
        const gb::RunStats s = gb::run_gpu(limit, segment_even, "", verbose != 0, start_even, sieve_split);
        fill(s, out);
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

}  // extern "C"
