/* goldbach_amd.h — run-level C ABI of the MI355X verifier.
 *
 * gba_run_gpu replaces the reference's host entry point
 *     RunStats gb::run_gpu(uint64_t limit, uint64_t segment_even,
 *                          const std::string& spv_path, bool verbose,
 *                          uint64_t start_even, uint32_t sieve_split_arg,
 *                          const char* prime_cache_path, bool collect_segments)
 * (include/verifier.hpp:103-110, src/verifier.cpp:2646) for callers that bind
 * C: plain integers in, a POD result out. Same range contract, same
 * fail-closed behaviour: any failure returns a non-zero status and nothing is
 * reported as PASS; the message is copied to `err`. The statistics files
 * (GB_STATS_STATE / GB_STATS_LOG / ...) are written exactly as by the C++
 * entry point.
 *
 * gba_segment_map exposes one segment's class map built on the GPU (the
 * reference's map validator role, tools/debug_map, which its tree does not
 * ship).
 */
#ifndef GOLDBACH_AMD_H
#define GOLDBACH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t limit, start, segments, leftovers;
    double init_seconds, gpu_seconds, total_seconds, evens_per_second;
    double sieve_seconds, large_seconds, verify_seconds, count_seconds, sync_seconds;
    uint64_t total_tiles, total_large_items;
    uint32_t split_used;
    int32_t stats_on;
    uint64_t stats_pmax, stats_primes, stats_twins, stats_sampled, stats_tail;
    /* device time of the hit-list harvest and window sieve kernels alone (HIP
     * events around each launch on the run's stream), their launches, and
     * their algorithmic HBM bytes over the run */
    double harvest_kernel_seconds, window_kernel_seconds;
    uint64_t harvest_launches, window_launches;
    double harvest_bytes, window_bytes;
    double host_plan_seconds;
    uint64_t harvest_reruns;       /* batches re-run with counted lists after a harvest region overflow */
    uint64_t graph_launches;       /* batches launched as one captured HIP graph (GB_GRAPH) */
    double harvest_words;          /* of harvest_bytes: the 8-byte multiplier-bitmap words read (a model, as the bytes) */
} gba_run_result;

/* One segment of a run (the reference's SegmentStats, include/verifier.hpp:
 * 14-27), filled when collect_segments is set. */
typedef struct {
    uint64_t seg_start, seg_end, evens, tiles, large_items;
    uint32_t sieve_split, mask_count, leftovers, pad0;
    double sieve_seconds, large_seconds, check_seconds, offset_seconds, wall_seconds;
} gba_segment_stats;

/* The reference's eight arguments in its order (verifier.hpp:103-110):
 * spv_path is accepted and ignored (the device code is linked into this
 * library), prime_cache_path names the GBP2 base-prime cache stem (NULL:
 * none), collect_segments asks for one gba_segment_stats per segment: the
 * first seg_cap of them are written to segs and *n_segs receives the run's
 * segment count. Returns 0 on PASS. stats_line (if non-null, capacity
 * stats_cap) receives the run's GBSTATS line when statistics were on. */
int gba_run_gpu(uint64_t limit, uint64_t segment_even, const char* spv_path, int verbose, uint64_t start_even,
                uint32_t sieve_split, const char* prime_cache_path, int collect_segments, gba_run_result* out,
                gba_segment_stats* segs, size_t seg_cap, size_t* n_segs, char* stats_line, size_t stats_cap,
                char* err, size_t err_cap);

/* One segment's class map: words [8 * cw] (cw returned in *cw_out, at most
 * cap words are written), Q[8] first q per class, nc[8] bits per class.
 * Returns 0 on success. */
int gba_segment_map(uint64_t seg_start, uint64_t limit, uint64_t span, uint64_t* words, size_t cap,
                    uint32_t* cw_out, uint64_t* Q, uint32_t* nc, char* err, size_t err_cap);

/* Engine: the device state kept resident between runs (base primes up to
 * sqrt(max_limit), the prime tables in HBM, segment buffers for spans up to
 * max_span). No counterpart in the reference, which rebuilds it per process
 * (its init phases, verifier.cpp:2615-2642); for callers that verify many
 * ranges (serving, batches) without repeating the set-up. gba_engine_run has
 * the contract of gba_run_gpu with limit <= max_limit; one thread per
 * engine. */
typedef struct gba_engine gba_engine;
int gba_engine_open(uint64_t max_limit, uint64_t max_span, gba_engine** out, double* init_seconds, char* err,
                    size_t err_cap);
int gba_engine_run(gba_engine* e, uint64_t limit, uint64_t segment_even, uint64_t start_even, gba_run_result* out,
                   char* stats_line, size_t stats_cap, char* err, size_t err_cap);
void gba_engine_close(gba_engine* e);

/* Builds, ahead of the first run, the engine's resident tables that depend
 * on the range (the multiplier bitmap of the large primes) for runs over
 * [start_even, limit]; a run builds them itself when they are missing.
 * *seconds receives the build time (0 when already resident). Counterpart of
 * the reference's start-up table build (build_mtab, src/verifier.cpp:
 * 1444-1508). Returns 0 on success. */
int gba_engine_prepare(gba_engine* e, uint64_t limit, uint64_t start_even, double* seconds, char* err,
                       size_t err_cap);

/* Batcher: many ranges through one engine's pipeline (batched and served
 * verification; no counterpart in the reference, which verifies one range
 * per process). A batch packs waiting segments of several ranges (each range
 * segmented exactly as a plain run of it), one launch per stage; every
 * range's statistics are those of a plain run of that range alone, without
 * campaign files (gba_batcher_result returns its GBSTATS line and the state
 * text a fresh campaign would hold after it).
 *   open:   plans for ranges with start >= min_start, limit <= max_limit
 *           (large-prime split, multiplier bitmap, buffers; the engine must
 *           cover max_limit); *batch_segments receives the segments a batch
 *           may hold. The batch buffers are the batcher's own, allocated
 *           here; a nonzero device_budget (bytes) bounds them: the segments
 *           per batch are the most whose buffers fit in 7/8 of it, and the
 *           direct harvest's regions are allocated only where they fit in
 *           the rest (else the counted form). Fails when one-segment batches
 *           do not fit.
 *   memory: the device bytes the batch buffers hold now, and the segments
 *           that can be in flight at once (two batches).
 *   segments: the segments [start_even, limit] splits into (as a plain run
 *           splits it); an admission policy reserves in-flight segments by it.
 *   submit: queues [start_even, limit], *id = 0, 1, ...
 *   step:   issues the next batch of waiting segments and collects the one
 *           issued before; returns up to finished_cap ids of completed
 *           ranges (the rest on later calls); *more = 0 when nothing waits,
 *           is in flight or is left to report.
 *   totals: kernel seconds, launches and algorithmic bytes over all batches.
 * A failed step (device error, failed check) leaves the batcher unusable.
 * One thread per batcher; the engine must outlive it. All return 0 on
 * success. */
typedef struct gba_batcher gba_batcher;
int gba_batcher_open(gba_engine* e, uint64_t max_limit, uint64_t min_start, uint64_t segment_even,
                     uint64_t device_budget, gba_batcher** out, uint32_t* batch_segments, char* err, size_t err_cap);
int gba_batcher_submit(gba_batcher* b, uint64_t start_even, uint64_t limit, uint64_t* id, char* err, size_t err_cap);
int gba_batcher_step(gba_batcher* b, uint64_t* finished, uint32_t finished_cap, uint32_t* n_finished, int* more,
                     char* err, size_t err_cap);
int gba_batcher_result(gba_batcher* b, uint64_t id, gba_run_result* out, char* stats_line, size_t stats_cap,
                       char* state, size_t state_cap, char* err, size_t err_cap);
int gba_batcher_totals(gba_batcher* b, gba_run_result* out);
int gba_batcher_memory(gba_batcher* b, uint64_t* device_bytes, uint32_t* inflight_segments);
int gba_batcher_segments(gba_batcher* b, uint64_t start_even, uint64_t limit, uint64_t* segments, char* err,
                         size_t err_cap);
void gba_batcher_close(gba_batcher* b);

/* Tuned knob defaults (the autotuner's file, goldbach-l0_amd/autotune.py):
 * installs the knobs of the first band of `path` that holds `limit` as the
 * defaults of knobs the environment does not set (path NULL: drops them);
 * `applied` receives "KNOB=value ...". An engine does this itself for its
 * limit when GB_TUNE_FILE is set. No counterpart in the reference (its
 * knobs are environment variables only). Returns 0 on success. */
int gba_tune_load(const char* path, uint64_t limit, char* applied, size_t applied_cap, char* err, size_t err_cap);

/* The library's version string. */
const char* gba_version(void);

#ifdef __cplusplus
}
#endif
#endif
