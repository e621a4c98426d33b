// GPU verification driver (contracts in verifier.hpp). The run is the
// reference's DenseRun (src/verifier.cpp:2615-2642) restated for the MI355X
// pipeline: range resolution and segmentation exactly as the reference
// (statistics depend on them), then batches of consecutive segments, each
// batch one launch per stage on one in-order HIP stream while the host
// collects the previous batch (two output slots), fail-closed coverage
// accounting, the leftover CPU fallback and the exact statistics.
#include "verifier.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <optional>
#include <stdexcept>
#include <thread>

#include "config.hpp"
#include "env.hpp"
#include "gbh.h"
#include "geometry.hpp"
#include "gstats.hpp"
#include "mbits.hpp"
#include "primes.hpp"
#include "ridge.hpp"
#include "tune.hpp"

namespace gb {
namespace {

using clock_ = std::chrono::steady_clock;
double since(clock_::time_point t0) { return std::chrono::duration<double>(clock_::now() - t0).count(); }

void check(int rc, const char* what)
{
    if (rc) throw std::runtime_error(std::string(what) + ": " + gbh_last_error());
}

struct Ctx {
    gbh_ctx* c = nullptr;
    Ctx() { check(gbh_open(0, &c), "gbh_open"); }
    ~Ctx() { gbh_close(c); }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
};

struct Buf {
    gbh_ctx* c = nullptr;
    void* p = nullptr;
    int kind = 0;
    size_t bytes = 0;
    uint64_t gen = 0;                  // bumped by every allocation (a new buffer's contents are unknown)
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { release(); }
    void release()
    {
        if (p) (void)gbh_free(c, p, kind);
        p = nullptr;
        bytes = 0;
    }
    // at least n bytes (contents are not kept on growth)
    void ensure(gbh_ctx* ctx, size_t n, int k)
    {
        n = std::max<size_t>(n, 256);
        if (p && bytes >= n) return;
        release();
        c = ctx;
        kind = k;
        check(gbh_alloc(c, n, k, &p), "gbh_alloc");
        bytes = n;
        ++gen;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Event {
    gbh_ctx* c = nullptr;
    gbh_event* e = nullptr;
    Event() = default;
    Event(const Event&) = delete;
    Event& operator=(const Event&) = delete;
    void create(gbh_ctx* ctx) { c = ctx; check(gbh_event_create(c, &e), "gbh_event_create"); }
    ~Event() { if (e) gbh_event_destroy(c, e); }
};

struct Graph {
    gbh_ctx* c = nullptr;
    gbh_graph* g = nullptr;
    Graph() = default;
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    void create(gbh_ctx* ctx) { c = ctx; check(gbh_graph_create(c, &g), "gbh_graph_create"); }
    ~Graph() { if (g) gbh_graph_destroy(c, g); }
};

double elapsed_s(gbh_ctx* c, const Event& a, const Event& b)
{
    float ms = 0;
    check(gbh_event_elapsed_ms(c, a.e, b.e, &ms), "gbh_event_elapsed_ms");
    return ms * 1e-3;
}

unsigned host_threads()
{
    const unsigned h = std::thread::hardware_concurrency();
    return std::max(1u, std::min(h ? h : 1u, 16u));
}

// Candidate primes 3..9781 and, per even class, the device's candidate slot
// -> prime index map (the useful candidates ascending, as gbh_api.hip builds
// them).
struct Candidates {
    std::vector<uint32_t> primes;
    std::vector<uint16_t> slot_pi[15];
    Candidates()
    {
        primes = odd_primes_up_to(MAX_GB_PRIME);
        for (uint32_t cr = 0; cr < 15; ++cr)
            for (size_t i = 0; i < primes.size(); ++i) {
                const uint32_t res = (2 * cr + 30 - primes[i] % 30) % 30;
                if (res % 2 && res % 3 && res % 5) slot_pi[cr].push_back(uint16_t(i));
            }
    }
};

// Batch events: stream 0 carries the uploads, the harvest, the hit-list
// phase of the window stage, the checker and the counter; stream 1 the
// set-up and the sieve phase of the window stage, beside the harvest.
// A batch's host-built tables, uploaded in one copy: the geometry of its
// segments, its runs and its counter parameters, each section 256-byte
// aligned.
struct TabLayout {
    size_t runs, cs, bytes;
};
inline TabLayout tab_layout(size_t nseg)
{
    const auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    TabLayout t;
    t.runs = al(sizeof(gbh_geom) * nseg);
    t.cs = al(t.runs + sizeof(gbh_hl_run) * nseg);
    t.bytes = t.cs + sizeof(gbh_countseg) * nseg;
    return t;
}

enum { EV_BEGIN, EV_HL0, EV_HL1, EV_SETUP, EV_WINDOW, EV_CHECK, EV_COUNT, EV_END, EV_GEOM, EV_S0, EV_S1, EV_COFF, EV_CA,
       EV_N };

// Knobs of removed reference pipelines are refused, not ignored: a run that
// asks for one would otherwise silently run something else
// (src/verifier.cpp:577-589 of the reference).
void refuse_removed_knobs()
{
    for (const char* gone : {"GB_SPLIT", "GB_SPLIT_CAP", "GB_SPLIT_STAGES", "GB_TWIN", "GB_TWIN_CAP", "GB_C30",
                             "GB_CLASSMAP", "GB_HITLIST", "GB_MASK_GROUPS", "GB_C30_LATE", "GB_C30_PREB",
                             "GB_ML_OVERLAP", "GB_HL_LAM2", "GB_HL_CAPG", "GB_HLH_WG", "GB_HL_UNCOMP", "GB_MT_UNCOMP",
                             "GB_MASK_XP", "GB_MED_ITEMS", "GB_MINP_FLOOR", "GB_HL_FILLDBG"})
        if (env::present(gone))
            throw std::runtime_error(std::string(gone) + " is a removed knob (the mod-30 map and checker are the only "
                                     "pipeline)");
}

// wheel fraction of the harvest: multipliers coprime to 7, 11 and 13
constexpr double kWheelFrac = (6.0 / 7) * (10.0 / 11) * (12.0 / 13);

struct Range;

// One segment of a batch, host side.
struct SegJob {
    Range* r = nullptr;          // the range it belongs to
    SegmentGeometry sg;
    ClassGeo cg;
    uint64_t index = 0;          // segment index within its range
    uint64_t n_gpu = 0;
    uint64_t gpu_evens = 0;
    bool sampled = false, walked = false, counted = false;
};

// Harvest cost of one run of a batch: its large primes, class bits, first q
// and hit density.
struct RunCost {
    uint32_t np = 0;
    double bits = 0, q_lo = 3, density = 0;
};

bool close_rel(double a, double b) { return std::fabs(a - b) <= 1e-3 * std::max(std::fabs(a), std::fabs(b)); }

// Runs of the same shape: the same prime counts, class bits, first q and
// hit density within 0.1% (a chunk plan balances load; any plan is exact).
bool same_shape(const std::vector<RunCost>& a, const std::vector<RunCost>& b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].np != b[i].np || !close_rel(a[i].bits, b[i].bits) || !close_rel(a[i].q_lo, b[i].q_lo) ||
            !close_rel(a[i].density, b[i].density))
            return false;
    return true;
}

// The harvest chunk plan last uploaded to an engine's device tables (chunk
// bounds, direct-harvest regions and partition groups): a batch of the same
// shape (the next batch of a run, the next run of a campaign) reuses it.
struct ChunkPlan {
    bool valid = false;
    std::vector<RunCost> key;
    uint32_t nbb = 0, nwheel = 0, nwheel_hi = 0, nwheel_hi2 = 0, nrand = 0, nx = 0;
    double ncb = 0;
    std::string mode;
    uint32_t region_fix = 0, chunks_cu = 0;
    uint32_t nfew = UINT32_MAX, nsingle = UINT32_MAX;
    std::vector<uint32_t> chunks, rdesc, groups;   // the tables uploaded to a slot
    uint64_t region_entries = 0;
    uint32_t nchunks = 0, stage = 0, ngroups = 0, part_cap = GBH_PART_CAP;
    bool direct = false;
};

// The device buffers of the batches in flight: two sets, so a batch's sieve
// and harvest run while the batch before it is checked and counted. A plain
// run uses its engine's set; a Batcher owns a set of its own, sized within
// its device budget (0: no budget).
struct Slots {
    struct Slot {
        Buf map, i0, coff, coffa, pending, scratch, hits, hit_n, chunks, rdesc, groups, region, fills;
        Buf ridge_more;                     // ridge records beyond the output block's GBH_RIDGE_CAP
        Buf hq;                             // the persistent harvest's chunk counter
        Buf tab;                            // the batch tables (tab_layout)
        Buf vmap, vhits, vhit_n, vchunks;   // the GB_BUCKET_VERIFY rebuild
    } sl[2];
    Buf h_vchunks;
    // two device output blocks per slot, used by its batches in turn: the
    // block the next batch takes is cleared while the batch before it runs
    Buf d_out[2][2];
    uint32_t out_sel[2] = {0, 0};             // the block the slot's next batch takes
    Buf h_out[2], h_tab[2], h_chunks[2], h_rdesc[2];
    Event ev[2][EV_N];
    // captured batches (graph mode): the events above only order the streams
    // inside the graph; tbeg times a captured batch from before its launch
    // (to ev[EV_END]); one graph per slot, its executable form updated in
    // place batch after batch
    Event tbeg[2];
    Graph graph[2];
    uint64_t slot_plan[2] = {~0ull, ~0ull};   // the chunk plan each slot's device tables hold
    uint64_t budget = 0;                      // device bytes the set may hold (0: unbounded)
    // output block k of a slot (its first clean_n segments) and its harvest
    // chunk counter (word 32 k of the slot's hq) are zero, cleared in stream
    // order by the slot's batch before, while the buffers are the
    // allocations clean_gen names: the batch that takes the block then need
    // not wait for the clears
    uint32_t clean_n[2][2] = {};
    uint64_t clean_gen[2][2][2] = {};

    void init(gbh_ctx* c, uint32_t kbatch)
    {
        for (int s = 0; s < 2; ++s) {
            for (Buf& o : d_out[s]) o.ensure(c, sizeof(gbh_segout) * kbatch, 0);
            h_out[s].ensure(c, sizeof(gbh_segout) * kbatch, 1);
            h_tab[s].ensure(c, tab_layout(kbatch).bytes, 1);
            for (auto& e : ev[s]) e.create(c);
            tbeg[s].create(c);
            graph[s].create(c);
        }
        for (auto& B : sl) {
            B.tab.ensure(c, tab_layout(kbatch).bytes, 0);
            B.coff.ensure(c, size_t(kbatch) * 15 * GBH_NCAND * 4, 0);
            B.coffa.ensure(c, size_t(kbatch) * 15 * GBH_NCAND_PAD * 4, 0);
        }
    }
    /// Device bytes held now.
    uint64_t device_bytes() const
    {
        uint64_t n = 0;
        for (const Slot& B : sl)
            for (const Buf* b : {&B.map, &B.i0, &B.coff, &B.coffa, &B.pending, &B.scratch, &B.hits, &B.hit_n,
                                 &B.chunks, &B.rdesc, &B.groups, &B.region, &B.fills, &B.ridge_more, &B.hq, &B.tab,
                                 &B.vmap, &B.vhits, &B.vhit_n, &B.vchunks})
                n += b->kind == 0 ? b->bytes : 0;
        for (const auto& pair : d_out)
            for (const Buf& b : pair) n += b.bytes;
        return n;
    }
};

}  // namespace

// ---------------------------------------------------------------- Engine state

struct Engine::State {
    clock_::time_point t0;
    uint64_t max_limit = 0, max_span = 0;
    std::vector<uint32_t> base;          // odd primes <= sqrt(max_limit) + 16
    uint32_t win_lo = 0, win_hi = 0;     // window primes base[win_lo, win_hi), large base[win_hi, ...)
    uint32_t ka = 0, kbatch = 0, win_kb = 0;
    double init_s = 0, cache_s = 0, module_s = 0, upload_s = 0;
    std::vector<double> inv_coarse;      // sum of 1/p over the first 256 k large primes
    /// sum of 1/p over the first j large primes
    double inv_at(uint32_t j) const
    {
        const uint32_t k = j >> 8;
        double s = inv_coarse[k];
        for (uint32_t i = k << 8; i < j; ++i) s += 1.0 / base[win_hi + i];
        return s;
    }
    std::optional<Ctx> ctx;
    gbh_ctx* c = nullptr;
    int cus = 256;
    Buf d_win;                           // window primes, then the large primes below 2^21 (one table)
    Buf d_pbase, d_pgap;                 // every large prime as the packed stream (gbh_pstream)
    gbh_pstream large{};                 // the harvest's view of it
    uint64_t win_p_long = 0;             // long runs' window bound (GB_WIN_P_LONG)
    Slots own;                           // the batch buffers of plain runs
    Candidates cand;
    ChunkPlan cplan;                     // the last harvest chunk plan made (host side)
    uint64_t cplan_version = 0;          // bumped by every new chunk plan
    std::string tuned;                   // knob defaults taken from GB_TUNE_FILE

    State(uint64_t max_limit_, uint64_t max_span_, const char* prime_cache)
    {
        t0 = clock_::now();
        require_supported_limit(max_limit_);
        // the autotuner's knob defaults for this engine's range (before any knob is read)
        if (const char* tf = std::getenv("GB_TUNE_FILE")) tuned = tune_load(tf, max_limit_);
        else tune_clear();
        max_limit = max_limit_;
        max_span = std::max<uint64_t>(max_span_ + (max_span_ & 1), 4);
        if (max_span > (uint64_t(1) << 35)) throw std::runtime_error("segment span above 2^35 integers");
        win_kb = env::u32("GB_WIN_KB", 0, 0, 128);
        if (win_kb & (win_kb - 1)) throw std::runtime_error("GB_WIN_KB must be a power of two");
        ka = env::u32("GB_KA", KA_DEFAULT, 4, GBH_KA_MAX);
        if (ka % 16) throw std::runtime_error("GB_KA must be a multiple of 16");
        kbatch = env::u32("GB_BATCH", 16, 1, GBH_MAX_BATCH);
        // base primes (the reference loads them from its prime cache;
        // here a segmented sieve on the host threads)
        const auto tc = clock_::now();
        base = base_primes_cached(std::max<uint32_t>(base_prime_bound(max_limit), 64), prime_cache, host_threads());
        cache_s = since(tc);
        win_lo = uint32_t(std::lower_bound(base.begin(), base.end(), 67u) - base.begin());
        // within [2^16, 2^21]: the hit lists' double-precision quotients q / p
        // stay within one of the integer quotient only for p >= 2^16, the
        // window's float remainders are exact for p < 2^21
        const uint64_t win_p = env::u64("GB_WIN_P", WIN_P_DEFAULT, uint64_t(1) << 16, uint64_t(1) << 21);
        win_p_long = env::u64("GB_WIN_P_LONG", WIN_P_LONG_DEFAULT, win_p, uint64_t(1) << 21);
        win_hi = std::max(win_lo, uint32_t(std::lower_bound(base.begin(), base.end(), win_p) - base.begin()));
        {
            const size_t nl = base.size() - win_hi;
            inv_coarse.assign(nl / 256 + 2, 0.0);
            double acc = 0;
            for (size_t i = 0; i < nl; ++i) {
                if (i % 256 == 0) inv_coarse[i / 256] = acc;
                acc += 1.0 / base[win_hi + i];
            }
            for (size_t k = (nl + 255) / 256; k < inv_coarse.size(); ++k) inv_coarse[k] = acc;
        }
        const auto tm = clock_::now();
        ctx.emplace();
        c = ctx->c;
        gbh_device_info_t info{};
        check(gbh_device_info(c, &info), "gbh_device_info");
        cus = info.cus > 0 ? info.cus : 256;
        module_s = since(tm);
        const auto tu = clock_::now();
        const uint32_t nwin = win_hi - win_lo, nlarge = uint32_t(base.size()) - win_hi;
        // the window stage reads the window primes and, on long runs, the
        // large primes below GB_WIN_P_LONG (<= 2^21) from one u32 table; the
        // harvest reads the large primes only from the packed stream
        const uint32_t nlow = uint32_t(std::lower_bound(base.begin() + win_hi, base.end(), uint32_t(1) << 21) -
                                       (base.begin() + win_hi));
        d_win.ensure(c, size_t(nwin + nlow) * 4 + 4, 0);
        if (nwin + nlow)
            check(gbh_copy_h2d(c, 0, d_win.p, base.data() + win_lo, size_t(nwin + nlow) * 4), "upload primes");
        if (nlarge) {
            PackedPrimes pk = pack_primes(base.data() + win_hi, nlarge, GBH_PS_GROUP, GBH_PS_PAD);
            d_pbase.ensure(c, pk.base.size() * 4, 0);
            d_pgap.ensure(c, pk.gap.size(), 0);
            check(gbh_copy_h2d(c, 0, d_pbase.p, pk.base.data(), pk.base.size() * 4), "upload packed primes");
            check(gbh_copy_h2d(c, 0, d_pgap.p, pk.gap.data(), pk.gap.size()), "upload packed primes");
            check(gbh_stream_sync(c, 0), "upload sync");   // the host vectors are freed next
            large = gbh_pstream{d_pbase.as<uint32_t>(), static_cast<const uint8_t*>(d_pgap.p), nlarge, 0};
        }
        own.init(c, kbatch);
        check(gbh_stream_sync(c, 0), "upload sync");
        upload_s = since(tu);
        init_s = since(t0);
    }
};

namespace {

// One range of a pipeline: the reference's DenseRun over [start, limit]
// (src/verifier.cpp:2615-2642): its resolved segmentation, statistics and
// ridge set-up, accounting and segment cursor. A plain run is one range; a
// batched or served workload is many, sharing the pipeline's batches.
struct Range {
    size_t id = 0;
    uint64_t start = 4, limit = 0, span = 0, total_evens = 0;
    uint64_t q_lo = 3;                 // the first q of its first segment
    uint32_t nwin = 0, nlarge = 0, split = 0;
    bool campaign = true;              // statistics and ridge files (plain runs)
    bool verbose = false, collect = false;
    bool stats_on = false, gaps_on = false, gaps_exact = false;
    uint32_t sample = 256;
    std::optional<GoldbachStats> gs;
    std::optional<RidgeTracker> ridges;
    std::string ridge_path;
    uint64_t sampled_total = 0, evens_done = 0;
    double density = 0;                // expected harvest hits per class bit over its large primes
    double inv_large = 0;              // sum of 1/p over its large primes
    RunStats rs;
    uint64_t next_start = 0, next_index = 0;
    bool issued_all = false, last_collected = false, finished = false, started = false;
    clock_::time_point t_submit, t_first;
    double t_sieve = 0, t_large = 0, t_check = 0, t_count = 0;
    std::string state_text;            // without campaign files: the state a fresh campaign holds after it
};

// Range resolution (the reference's resolve_range): even bounds, the span
// clamp to at least two segments, the run's window / large prime counts.
void resolve_range(Engine::State& E, Range& r, uint32_t split_arg)
{
    refuse_removed_knobs();
    require_supported_limit(r.limit);
    r.limit &= ~uint64_t(1);
    r.rs.limit = r.limit;
    if (r.limit < 4) throw std::runtime_error("limit < 4: no even n >= 4 to verify");
    if (r.limit > E.max_limit) throw std::runtime_error("limit above the engine's prime bound");
    if (r.start < 4) r.start = 4;
    if (r.start & 1) r.start += 1;
    if (r.start > r.limit) throw std::runtime_error("no even n in [start, limit]");
    uint64_t& span = r.span;
    if (span < 2) span = 2;
    if (span & 1) span += 1;
    r.total_evens = (r.limit - r.start) / 2 + 1;
    // at least two segments (the span is an integer span: span / 2 evens)
    if (span > r.total_evens) span = std::max<uint64_t>(r.total_evens & ~uint64_t(1), 2);
    if (r.start == 4) {
        if (r.limit < 6)
            throw std::runtime_error("the GPU path needs limit >= 6 (n = 4 = 2 + 2 has no odd-prime partition; "
                                     "--cpu-check 4 verifies it)");
        span = std::max<uint64_t>(span, 4);
    }
    if (span > E.max_span) throw std::runtime_error("segment span above the engine's buffers");
    // a segment's prime positions (q - q_first, the counter's gap walk) are
    // 32-bit: the merged last segment (below two spans plus the 9781 below
    // its first even) must stay below 2^32 integers
    if (span >= (uint64_t(1) << 31) - (uint64_t(1) << 20))
        throw std::runtime_error("segment span of 2^31 - 2^20 integers or more (32-bit prime positions)");
    // this run's primes: those <= sqrt(limit) + 16
    const uint32_t bound = std::max<uint32_t>(base_prime_bound(r.limit), 64);
    const uint32_t nb = uint32_t(std::upper_bound(E.base.begin(), E.base.end(), bound) - E.base.begin());
    uint32_t hi = std::min(E.win_hi, nb);
    if (split_arg > 0) {          // --split: the first large prime's index among the base primes
        if (split_arg < E.win_lo || split_arg > E.win_hi)
            throw std::runtime_error("--split must lie in [" + std::to_string(E.win_lo) + ", " +
                                     std::to_string(E.win_hi) + "] on this engine");
        if (split_arg < hi)
            std::fprintf(stderr, "note: --split below the engine's window bound: the windows keep their primes\n");
    }
    r.nwin = hi > E.win_lo ? hi - E.win_lo : 0;
    r.split = E.win_lo + r.nwin;
    r.nlarge = nb > E.win_hi ? nb - E.win_hi : 0;
    r.rs.split_used = r.split;
    r.q_lo = r.start >= MAX_GB_PRIME + 2 ? r.start - MAX_GB_PRIME : 3;
    r.inv_large = r.nlarge ? E.inv_at(r.nlarge) : 0.0;
    r.next_start = r.start;
}

// Goldbach ridges (the reference's configure_ridges_and_stats,
// src/verifier.cpp:718-744): the floor is clamped above every phase-A
// candidate, so above it phase B sees every even; a campaign run loads the
// store after the clamp.
void configure_ridges(Engine::State& E, Range& r)
{
    r.ridges.emplace(&E.cand.primes);
    uint32_t pa_max = 0;
    for (uint32_t cr = 0; cr < 15; ++cr)
        pa_max = std::max(pa_max, E.cand.primes[E.cand.slot_pi[cr][std::min<size_t>(E.ka, E.cand.slot_pi[cr].size()) - 1]]);
    if (r.ridges->active()) {
        r.ridges->clamp_floor_above(pa_max);
        if (r.campaign) {
            r.ridge_path = env::path("GB_RIDGE_FILE", "goldbach_ridges.csv");
            r.ridges->load(r.ridge_path);
        }
    }
}

void configure_stats(Engine::State& E, Range& r)
{
    const bool stats_env = env::onoff("GB_STATS", true);
    r.stats_on = r.start >= STATS_MIN_START && stats_env;
    const std::string gaps_env = env::choice("GB_GAPS", "exact", {"exact", "1", "sampled", "0"});
    r.gaps_on = r.stats_on && gaps_env != "0";
    r.gaps_exact = r.gaps_on && gaps_env != "sampled";
    r.sample = env::u32("GB_STATS_SAMPLE", 256, 1);
    if (r.stats_on) {
        r.gs.emplace(E.cand.primes, CHECK_TAIL, r.campaign);
        r.gs->set_run_start(r.start);
        if (r.gaps_on) r.gs->begin_gaps(r.start);
    } else if (r.campaign && r.start < STATS_MIN_START && !env::present("GB_STATS")) {
        std::fprintf(stderr, "note: Goldbach statistics off below start %llu\n", (unsigned long long)STATS_MIN_START);
    }
    // every phase-A candidate must lie below the tail bound: evens with
    // p(n) >= tail_t then all reach phase B, which records them exactly
    for (uint32_t cr = 0; cr < 15; ++cr)
        if (E.cand.primes[E.cand.slot_pi[cr][std::min<size_t>(E.ka, E.cand.slot_pi[cr].size()) - 1]] >= CHECK_TAIL)
            throw std::runtime_error("GB_KA too large: a phase-A candidate reaches the tail bound");
}

// What a pipeline is planned for: the largest q and its prime counts, the
// smallest q whose range has large primes, the largest span, the segments a
// plain run's batches hold, and the first q of the ranges whose hit density
// sizes the lists.
struct PlanBounds {
    uint64_t q_hi = 0;
    uint64_t qspan = 0;               // q span of a batch's run (0: a full batch of spans)
    uint64_t sq_lo = 0;               // floor(sqrt(smallest q_lo with large primes)); 0: no large primes
    uint64_t span = 0;
    uint32_t nwin = 0, nlarge = 0;
    uint64_t nseg = 1;
    std::vector<std::pair<uint64_t, uint32_t>> q_lo_np;   // (q_lo, large primes) of the ranges to serve
};

PlanBounds bounds_of(const Range& r)
{
    PlanBounds b;
    b.q_hi = r.limit - 3;
    b.qspan = r.limit - r.start + 2;
    b.sq_lo = r.nlarge ? std::max<uint64_t>(uint64_t(std::sqrt(double(r.q_lo))), 1) : 0;
    b.span = r.span;
    b.nwin = r.nwin;
    b.nlarge = r.nlarge;
    b.nseg = (r.total_evens + r.span / 2 - 1) / std::max<uint64_t>(r.span / 2, 1) + 1;
    b.q_lo_np.push_back({r.q_lo, r.nlarge});
    return b;
}


// The segment pipeline of one engine: batches of segments (of one range, or
// of several ranges as runs), each batch one launch per stage on one
// in-order HIP stream while the host collects the previous batch (two
// output slots), fail-closed coverage accounting, the leftover CPU fallback
// and the exact statistics of every range.
class Pipeline {
public:
    Pipeline(Engine::State& e, bool multi, Slots* slots = nullptr) : E(e), S(slots ? *slots : e.own), multi_(multi)
    {
        phase_trace_ = env::present("GB_PHASE");
        if (kend_ && kend_ <= E.ka) throw std::runtime_error("GB_PHASEB_SLOTS must lie above GB_KA (0: every slot)");
        graph_ = env::onoff("GB_GRAPH", false);
    }
    /// The large primes' split and the multiplier bitmap for @p b.
    void plan_bitmap(const PlanBounds& b);
    /// Bitmap, window size, list capacities and buffers for @p b.
    void plan(const PlanBounds& b);
    double density_of(uint64_t q_lo, uint32_t nlarge) const;
    void add(Range* r) { active_.push_back(r); }
    /// Issues the next batch of waiting segments and collects the batch
    /// before it; ranges whose last segment was collected are finished and
    /// appended to @p done. @return whether anything is waiting or in flight.
    bool step(std::vector<Range*>* done);
    /// The class map of one segment of @p r alone (validators).
    uint32_t map_of(Range& r, uint64_t seg_start, std::vector<uint64_t>& words, uint64_t Q[8], uint32_t nc[8]);
    uint32_t kb() const { return kb_; }
    const RunStats& totals() const { return agg_; }
    const Slots& slots() const { return S; }

private:
    double hit_factor(uint32_t j, double q_lo) const;
    void plan_chunks(const std::vector<RunCost>& runs, uint32_t nbb, double ncb, int slot);
    void upload_chunks(int slot);
    uint32_t run_cap(const Range& r) const;
    void form(std::vector<SegJob>& bt);
    void enqueue(std::vector<SegJob>& batch, int slot, bool check_and_count);
    void enqueue_verify(uint32_t s, const gbh_batch& b, gbh_segout* out, int slot);
    void collect(std::vector<SegJob>& batch, int slot);
    void collect_segment(const SegJob& j, const gbh_segout& o, const double share[4], int slot, uint32_t s);
    void finish_range(Range& r);

    Engine::State& E;
    Slots& S;                          // the batch buffers this pipeline's batches use
    bool multi_;
    bool phase_trace_ = false;
    std::vector<Range*> active_;       // ranges with segments not yet issued, in submission order
    std::vector<SegJob> batches_[2];
    int cur_ = 0;
    bool pending_ = false;
    RunStats agg_;                     // kernel times and bytes over every batch
    uint32_t nwin_ = 0, nlarge_ = 0, ncoop_ = 0;   // the largest range's window / large prime counts
    // long runs: the first nx_ large primes (wheel primes below
    // GB_WIN_P_LONG) are sieved in the windows beside the window primes
    // (their multiples marked in LDS instead of listed); the harvest starts
    // at large prime nx_
    uint32_t nx_ = 0;
    uint32_t wbits_ = 0, nwin_max_ = 0, cw_ = 0, cap_ = 0, pcap_ = 0, swords_ = 0, kb_ = 1;
    uint64_t span_ = 0, evens_cap_ = UINT64_MAX;
    // large primes: [0, nwheel_) and the band [nwheel_hi_, nwheel_hi2_) walk the wheel, the others read
    // the multiplier bitmap (primes from t_hl_ on, except the band's t_hi_ .. below t_hi2_)
    uint32_t nwheel_ = 0, t_hl_ = UINT32_MAX;
    uint32_t nwheel_hi_ = UINT32_MAX, nwheel_hi2_ = UINT32_MAX, t_hi_ = UINT32_MAX, t_hi2_ = UINT32_MAX;
    double inv_hi_ = 0, inv_hi2_ = 0;   // sum of 1/p over the large primes below t_hi_, below t_hi2_
    bool on_wheel(uint32_t j) const { return j < nwheel_ || (j >= nwheel_hi_ && j < nwheel_hi2_); }
    // bitmap primes below nrand_ read their words at random (consecutive
    // primes' words more than a sector apart)
    uint32_t nrand_ = 0;
    bool random_reads(uint32_t j) const { return j >= nwheel_ && j < nrand_ && !on_wheel(j); }
    // bitmap primes from nfew_ on have at most four multipliers in a batch
    // run's q span, from nsingle_ on at most one (the harvest's compacted
    // path; set by the chunk plan)
    uint32_t nfew_ = UINT32_MAX, nsingle_ = UINT32_MAX;
    // the compacted path pays for its registers only when it serves a
    // quarter of the large primes or more
    bool few_path() const { return nfew_ < nlarge_ && 4ull * (nlarge_ - nfew_) >= nlarge_; }
    // a prime index rounded up to a group of the packed stream (at most nlarge_)
    uint32_t ps_up(uint32_t j) const
    {
        return std::min<uint32_t>(nlarge_, (j + GBH_PS_GROUP - 1) / GBH_PS_GROUP * GBH_PS_GROUP);
    }
    // the first chunk boundary above a: the wheel / bitmap / band /
    // random-read boundaries (chunks are all of one kind)
    uint32_t next_cut(uint32_t a, uint32_t end) const
    {
        for (uint32_t c : {nwheel_, nwheel_hi_, nwheel_hi2_, nrand_, nfew_, nsingle_})
            if (c > a && c < end) end = c;
        return end;
    }
    double inv_wheel_ = 0;             // sum of 1/p over the wheel primes
    double density_max_ = 0;
    std::shared_ptr<const MultBitmap> mb_;
    // harvest chunks of the batch in each slot (re-planned when the runs change)
    std::vector<uint32_t> chunks_;
    uint32_t nchunks_ = 0, stage_ = 0;
    // the harvest's form: LDS-staged (short batches), direct (one walk into
    // per-work-group regions, then the partition) or counted (two walks);
    // GB_HL_MODE = auto | direct | counted
    std::string hl_mode_ = env::choice("GB_HL_MODE", "auto", {"auto", "direct", "counted"});
    bool direct_ = false, force_count_ = false, direct_used_[2] = {false, false};
    bool count_side_[2] = {false, false};   // the slot's counter ran on its side stream
    // the window stage's sieve phase beside the harvest on a second stream (GB_OVERLAP=0: one stream)
    // GB_OVERLAP=0: one stream for everything (no stage or batch runs
    // beside another; for profiling and as a fallback)
    bool overlap_ = env::onoff("GB_OVERLAP", true);
    bool checkb_split_ = env::onoff("GB_CHECKB_SPLIT", true);
    // GB_GRAPH=1: each batch's launches captured as one HIP graph (the
    // slot's executable graph updated in place) and launched at once; 0
    // (default): issued one by one. Measured on MI355X with ROCm 7: graphs
    // 8% slower on a plain run (its harvest waits for the whole capture
    // instead of starting while the rest is queued) and 4% slower on the
    // batched ranges (whose next batch is captured while the GPU runs the
    // one before)
    bool graph_ = false;
    bool captured_[2] = {false, false};   // the slot's batch in flight ran as a graph
    /// Event `id` of the slot on `stream` (in a captured batch an ordering
    /// node of the graph, no timestamp).
    void mark(int stream, int slot, int id) { check(gbh_event_record(E.c, stream, S.ev[slot][id].e), "event"); }

    int main_stream(int slot) const { return overlap_ ? slot : 0; }
    uint64_t region_entries_ = 0;
    uint64_t planned_bytes_ = 0;       // the batch buffers' device bytes planned by plan()
    bool plan_copies_ = false;         // uploads issued since plan() last synchronised
    // GB_HL_HITQ: the direct harvest's bitmap walk files its hits through a
    // per-wave LDS queue (1) or where it finds them (0); the queue comes with
    // the compacted path's kernel instance (short batches), whose registers
    // it shares
    bool hit_queue_ = env::onoff("GB_HL_HITQ", true);
    uint32_t ngroups_ = 0, part_cap_ = GBH_PART_CAP;   // the partition: groups and their LDS entries
    // GB_BUCKET_VERIFY=N: every Nth segment's map is rebuilt with every large
    // prime on the wheel walk (no multiplier bitmap) and compared word for word
    uint32_t verify_every_ = 0, cap_v_ = 0, nvchunks_ = 0;
    // GB_PHASEB_SLOTS=K: phase B probes candidate slots below K only, so the
    // evens whose p(n) lies beyond become leftovers for the CPU (the role of
    // the reference's GB_C30_PLAN override: it exercises the fallback)
    uint32_t kend_ = env::u32("GB_PHASEB_SLOTS", 0, 0, GBH_NCAND);
    // GB_HL_REGION=N: every direct-harvest region holds N entries per list
    // instead of the size planned from the expected hits (a small N forces
    // the region overflow and the counted re-run of the batch)
    uint32_t region_fix_ = env::u32("GB_HL_REGION", 0, 0, GBH_PART_CAP);
    // GB_HL_PERSIST: the harvest as persistent work-groups taking chunks from
    // a device counter (1), or one work-group per chunk (0)
    bool persist_ = env::onoff("GB_HL_PERSIST", true);
    // GB_HL_PERSIST_WG: its work-groups per CU at most (0: as many as fit;
    // fewer leave room for the window sieve running beside it)
    uint32_t persist_wg_ = env::u32("GB_HL_PERSIST_WG", 0, 0, 64);
    // GB_HL_CHUNKS_CU: harvest chunks per CU the chunk plan aims at (0:
    // 16 for batches of at most 32 (segment, class) lists, else 8: the
    // partition's work grows with the chunks times the lists)
    uint32_t chunks_cu_ = env::u32("GB_HL_CHUNKS_CU", 0, 0, 256);
    uint64_t verified_ = 0;
    bool harvested_[2] = {false, false};
};

// The large primes' split between the wheel walk and the multiplier bitmap:
// a prime p >= t_hl lists p m with m prime, exact while every q < t_hl^3;
// the bitmap covers m in [max(t_hl, sqrt(q_lo)), q_hi / t_hl] within its
// budget (t_hl rises until it fits). One split serves every range of the
// pipeline (each range's multipliers lie inside the bitmap's span).
void Pipeline::plan_bitmap(const PlanBounds& b)
{
    nlarge_ = b.nlarge;
    nwheel_ = nlarge_;
    t_hl_ = UINT32_MAX;
    nwheel_hi_ = nwheel_hi2_ = nlarge_;
    t_hi_ = t_hi2_ = UINT32_MAX;
    nrand_ = 0;
    mb_.reset();
    if (!nlarge_ || !b.sq_lo) return;
    const uint64_t budget = mult_bitmap_budget();
    if (!budget) return;
    const uint64_t q_hi = b.q_hi, sq_lo = b.sq_lo;
    const uint64_t m_cap = (budget - std::min<uint64_t>(budget, uint64_t(1) << 24)) * 30;   // bitmap bytes -> m span
    const uint32_t* lp = E.base.data() + E.win_hi;
    // smallest bitmap prime: its cube above q_hi, and the m span within the budget
    const auto cube_le = [](uint64_t x, uint64_t q) { return (unsigned __int128)x * x * x <= q; };
    uint64_t cb = uint64_t(std::cbrt(double(q_hi)));
    while (cube_le(cb + 1, q_hi)) ++cb;
    while (cb > 1 && !cube_le(cb, q_hi)) --cb;
    uint64_t t = std::max<uint64_t>({lp[0], cb + 1, q_hi / (m_cap + sq_lo) + 1});
    // a resident bitmap serves this range too when t rises to its span
    // (m >= its base, q_hi / t below its end): taken when t rises at most
    // 8-fold (the primes in between walk the wheel instead)
    if (const auto cur = mult_bitmap_peek()) {
        const uint64_t m_end = cur->m_base + cur->m_span;   // exclusive
        uint64_t tc = std::max(t, q_hi / (m_end - 1) + 1);
        if (sq_lo < cur->m_base + 1) tc = std::max(tc, cur->m_base);
        if (tc <= 8 * t) t = tc;
    }
    nwheel_ = uint32_t(std::lower_bound(lp, lp + nlarge_, uint32_t(std::min<uint64_t>(t, UINT32_MAX))) - lp);
    // every boundary between prime kinds at a group of the packed stream
    // (harvest chunks start there): t rises to the next group's first prime
    nwheel_ = ps_up(nwheel_);
    if (nwheel_ == nlarge_) return;
    // a band of primes walks the wheel instead of reading the bitmap: those
    // whose walk lists fewer than GB_HL_WHEEL_HITS multiples per batch run
    // (0.192 qspan / p: the multipliers coprime to 30030) and whose bitmap
    // words are a random read each (consecutive primes' words further apart
    // than a 64-byte sector: p below sqrt(q ln p / 1920))
    {
        const double hits = double(env::u32("GB_HL_WHEEL_HITS", 2, 0, 1u << 20));
        const double qspan = double(b.qspan ? std::min<uint64_t>(b.qspan, uint64_t(E.kbatch) * b.span)
                                            : uint64_t(E.kbatch) * b.span);
        const double q = double(sq_lo) * double(sq_lo);
        const double th = hits > 0 ? 0.1918 * qspan / hits : 1e300;
        // ... and below a quarter of the span (from there the harvest's
        // compacted path scans them: at most four multipliers each)
        const double th2 = std::min(std::sqrt(q * std::log(std::sqrt(std::max(q, 4.0))) / 1920.0), qspan / 4);
        const auto at = [&](double x) {
            return ps_up(uint32_t(std::lower_bound(lp + nwheel_, lp + nlarge_,
                                                   uint32_t(std::min(std::max(x, 1.0), double(UINT32_MAX)))) - lp));
        };
        nrand_ = at(th2);
        if (th < th2 && th < double(UINT32_MAX)) {
            nwheel_hi_ = at(th);
            nwheel_hi2_ = at(th2);
            if (nwheel_hi2_ <= nwheel_hi_) {
                nwheel_hi_ = nwheel_hi2_ = nlarge_;
            } else {
                t_hi_ = lp[nwheel_hi_];
                t_hi2_ = nwheel_hi2_ < nlarge_ ? lp[nwheel_hi2_] : UINT32_MAX;
            }
        }
        if (nwheel_hi_ == nwheel_ && nwheel_hi2_ == nlarge_) {   // no prime left for the bitmap
            nwheel_ = nwheel_hi_ = nwheel_hi2_ = nlarge_;
            t_hi_ = t_hi2_ = UINT32_MAX;
            return;
        }
    }
    t_hl_ = lp[nwheel_];
    const uint64_t m_lo = std::max<uint64_t>(t_hl_, sq_lo > 1 ? sq_lo - 1 : 1), m_hi = q_hi / t_hl_;
    if (m_hi < m_lo) { nwheel_ = nwheel_hi_ = nwheel_hi2_ = nlarge_; t_hl_ = t_hi_ = t_hi2_ = UINT32_MAX; return; }
    mb_ = mult_bitmap(m_lo, m_hi, E.base, phase_trace_);
}

// Expected hits per class bit and unit 1/p of large prime j (the j-th above
// the window bound): the wheel lists every multiplier coprime to 7 11 13,
// the bitmap lists prime multipliers, (30/8) / ln m of those coprime to 30
// (m = q / p >= max(q_lo / p, p), the bound that makes the rate largest).
double Pipeline::hit_factor(uint32_t j, double q_lo) const
{
    if (on_wheel(j)) return kWheelFrac;
    const double p = E.base[E.win_hi + j];
    return 3.75 / std::log(std::max({q_lo / p, p, 3.0}));
}

// Expected hits per class bit over the first nlarge large primes of a range
// starting at q_lo: the wheel primes exactly, the bitmap primes in 64 bands
// of the prime index, each at its largest rate.
double Pipeline::density_of(uint64_t q_lo, uint32_t nlarge) const
{
    if (nlarge <= nx_) return 0.0;
    // every listed prime (from nx_ on) at the wheel's rate ...
    double d = kWheelFrac * (E.inv_at(nlarge) - E.inv_at(nx_));
    // ... except the bitmap's intervals, in 32 bands each at the band's largest rate
    const uint32_t iv[2][2] = {{std::min(nwheel_, nlarge), std::min(nwheel_hi_, nlarge)},
                               {std::min(nwheel_hi2_, nlarge), nlarge}};
    for (const auto& r : iv) {
        const uint32_t a0 = r[0], a1 = r[1];
        if (a1 <= a0) continue;
        double ia = E.inv_at(a0);
        const double i1 = E.inv_at(a1);
        d -= kWheelFrac * (i1 - ia);
        uint32_t a = a0;
        for (uint32_t k = 1; k <= 32 && a < a1; ++k) {   // band ends on the 1/p table's 256-prime grid
            const uint32_t b = k == 32 ? a1 : (a0 + uint32_t(uint64_t(a1 - a0) * k / 32)) & ~255u;
            if (b <= a) continue;
            const double ib = b == a1 ? i1 : E.inv_coarse[b >> 8];
            d += (ib - ia) * std::max(hit_factor(a, double(q_lo)), hit_factor(b - 1, double(q_lo)));
            a = b;
            ia = ib;
        }
    }
    return d;
}

// Window size, strides, hit-list capacity and buffer sizes.
void Pipeline::plan(const PlanBounds& pb)
{
    plan_bitmap(pb);
    nwin_ = pb.nwin;
    span_ = pb.span;
    const uint64_t nseg = std::min<uint64_t>(multi_ ? GBH_MAX_BATCH : E.kbatch, pb.nseg);
    nx_ = 0;
    if (!multi_ && nseg >= kLongBatch && nwin_ == E.win_hi - E.win_lo && nlarge_) {
        const uint32_t* lp = E.base.data() + E.win_hi;
        const uint32_t n = uint32_t(std::lower_bound(lp, lp + nlarge_, uint32_t(E.win_p_long)) - lp);
        nx_ = std::min(n, nwheel_) / GBH_PS_GROUP * GBH_PS_GROUP;   // the harvest's first chunk at a group
    }
    // the largest segment: the merged last one spans below 2 x span
    const uint64_t qodds_max = (2 * span_ + MAX_GB_PRIME) / 2 + 2;
    const uint64_t nc_max = (qodds_max + 14) / 15 + 1;
    const uint64_t nc_typ = (span_ / 2 + MAX_GB_PRIME / 2 + 14) / 15 + 1;
    // window: the largest power of two that still gives two work-groups per
    // CU over a batch of typical segments (GB_WIN_KB fixes it)
    if (E.win_kb) {
        wbits_ = E.win_kb * 8192;
    } else {
        wbits_ = 1u << 19;
        const uint64_t nseg_typ = multi_ ? E.kbatch : nseg;
        while (wbits_ > (1u << 15) && nseg_typ * 8 * ((nc_typ + wbits_ - 1) / wbits_) < 2ull * E.cus) wbits_ >>= 1;
    }
    nwin_max_ = uint32_t((nc_max + wbits_ - 1) / wbits_);
    while (nwin_max_ > 256 && wbits_ < (1u << 20)) {   // the window split's LDS tables
        wbits_ <<= 1;
        nwin_max_ = uint32_t((nc_max + wbits_ - 1) / wbits_);
    }
    if (nwin_max_ > 256) throw std::runtime_error("segment span too large for the window split");
    cw_ = nwin_max_ * (wbits_ / 64) + 16;
    ncoop_ = 0;
    while (ncoop_ < nwin_ && E.base[E.win_lo + ncoop_] < wbits_ / 128) ++ncoop_;
    inv_wheel_ = E.inv_at(nwheel_);
    inv_hi_ = E.inv_at(std::min(nwheel_hi_, nlarge_));
    inv_hi2_ = E.inv_at(std::min(nwheel_hi2_, nlarge_));
    density_max_ = 0;
    for (const auto& [q_lo, np] : pb.q_lo_np) density_max_ = std::max(density_max_, density_of(q_lo, np));
    // a batch's q range stays below 2^33 (the harvest's 32-bit half-offsets):
    // at most kb_ - 1 spans plus the merged last segment (below 2 spans)
    const uint64_t kmax = std::max<uint64_t>(1, ((uint64_t(1) << 33) - 4 * MAX_GB_PRIME) / span_ - 1);
    const double lam = wbits_ * density_max_;
    const uint64_t evens_max = span_ + 2;
    swords_ = uint32_t(2 * GBH_COUNT_SUB * (nc_max / 64 + 4));
    // list and pending capacities of batches of K segments
    auto caps = [&](uint64_t K) {
        return std::pair<uint32_t, uint32_t>{
            uint32_t(std::min(lam * 1.15 + 8 * std::sqrt(lam) + 1024, 4e9 / (double(K) * 8 * nwin_max_))),
            uint32_t(std::min<uint64_t>(evens_max / 32 + 65536, UINT32_MAX / 2 / K))};
    };
    // device bytes of the two slots' batch buffers for batches of K segments
    // (the ensure calls below)
    auto slot_bytes = [&](uint64_t K) {
        const auto [cap, pcap] = caps(K);
        auto b = [](uint64_t n) { return std::max<uint64_t>(n, 256); };
        const uint64_t one = b(K * 8 * cw_ * 8) + b(K * 8 * uint64_t(std::max(nwin_ + nx_, 1u)) * 4) +
                             b(K * uint64_t(pcap) * 4) + b(K * uint64_t(swords_) * 4) +
                             b(K * uint64_t(GBH_RIDGE_MORE) * 16) + 256 + b(K * 8 * uint64_t(nwin_max_) * 4) +
                             (nlarge_ ? b(K * 8 * uint64_t(nwin_max_) * cap * 4) : 0) + b(tab_layout(K).bytes) +
                             b(K * 15 * GBH_NCAND * 4) + b(K * 15 * GBH_NCAND_PAD * 4) + 2 * b(sizeof(gbh_segout) * K);
        return 2 * one;
    };
    if (multi_) {
        kb_ = GBH_MAX_BATCH;
        if (S.budget) {
            // the batch buffers within 7/8 of the budget; the rest holds the
            // chunk tables and the direct harvest's regions (else counted)
            const uint64_t room = S.budget / 8 * 7;
            while (kb_ > 1 && slot_bytes(kb_) > room) --kb_;
            if (slot_bytes(kb_) > room)
                throw std::runtime_error("device budget of " + std::to_string(S.budget) + " bytes below the buffers "
                                         "of one-segment batches (" + std::to_string(slot_bytes(1)) + " bytes, 7/8 "
                                         "of the budget at most)");
        }
        evens_cap_ = uint64_t(std::min(E.kbatch, kb_)) * (span_ / 2);
    } else {
        kb_ = uint32_t(std::min<uint64_t>({nseg, kmax, GBH_HL_BUCKETS / (8 * nwin_max_)}));
    }
    if (!kb_ || GBH_HL_BUCKETS / (8 * nwin_max_) == 0) throw std::runtime_error("segment span too large for the hit-list work-group");
    std::tie(cap_, pcap_) = caps(kb_);
    planned_bytes_ = slot_bytes(kb_);
    gbh_ctx* c = E.c;
    const size_t K = kb_;
    for (auto& B : S.sl) {
        B.map.ensure(c, K * 8 * cw_ * 8, 0);
        B.i0.ensure(c, K * 8 * size_t(std::max(nwin_ + nx_, 1u)) * 4, 0);
        B.pending.ensure(c, K * size_t(pcap_) * 4, 0);
        B.scratch.ensure(c, K * size_t(swords_) * 4, 0);
        B.ridge_more.ensure(c, K * size_t(GBH_RIDGE_MORE) * 16, 0);
        B.hq.ensure(c, 256, 0);
        B.hit_n.ensure(c, K * 8 * size_t(nwin_max_) * 4, 0);
        if (nlarge_) B.hits.ensure(c, K * 8 * size_t(nwin_max_) * cap_ * 4, 0);
        B.tab.ensure(c, tab_layout(K).bytes, 0);
        B.coff.ensure(c, K * 15 * GBH_NCAND * 4, 0);
        B.coffa.ensure(c, K * 15 * GBH_NCAND_PAD * 4, 0);
    }
    for (int s = 0; s < 2; ++s) {
        for (Buf& o : S.d_out[s]) o.ensure(c, sizeof(gbh_segout) * K, 0);
        S.h_out[s].ensure(c, sizeof(gbh_segout) * K, 1);
        S.h_tab[s].ensure(c, tab_layout(K).bytes, 1);
    }
    if (!multi_ && nlarge_) {   // a plain run's batches share one chunk plan: make it now
        const auto tp0 = clock_::now();
        plan_chunks({RunCost{nlarge_, double(kb_) * double(span_) * (8.0 / 30), double(pb.q_lo_np[0].first),
                             density_max_}}, kb_ * 8, double(nc_max), 0);
        if (phase_trace_) std::fprintf(stderr, "plan: chunking %.3f ms\n", since(tp0) * 1e3);
    }
    if (phase_trace_)
        std::fprintf(stderr, "plan: %u segments per batch, window %u bits, %u window primes, %u large (%u in the "
                     "windows, %u on the wheel, bitmap from %u, the wheel again in [%u, %u)), %u harvest chunks, list "
                     "cap %u, stage %u\n", kb_, wbits_, nwin_, nlarge_, nx_,
                     nwheel_ - std::min(nx_, nwheel_) + (std::min(nwheel_hi2_, nlarge_) - std::min(nwheel_hi_, nlarge_)),
                     t_hl_ == UINT32_MAX ? 0u : t_hl_, t_hi_ == UINT32_MAX ? 0u : t_hi_,
                     t_hi2_ == UINT32_MAX ? 0u : t_hi2_, nchunks_, cap_, stage_);
    verify_every_ = env::present("GB_BUCKET_VERIFY") ? env::u32("GB_BUCKET_VERIFY", 0, 0) : 0u;
    if (verify_every_ && nlarge_) {
        // one segment, every large prime on the wheel: its own lists and chunks
        const double inv_large = E.inv_at(nlarge_);
        const double lam_v = wbits_ * kWheelFrac * inv_large;
        cap_v_ = uint32_t(lam_v * 1.15 + 8 * std::sqrt(lam_v) + 1024);
        for (auto& B : S.sl) {
            B.vmap.ensure(c, size_t(8) * cw_ * 8, 0);
            B.vhits.ensure(c, size_t(8) * nwin_max_ * cap_v_ * 4, 0);
            B.vhit_n.ensure(c, size_t(8) * nwin_max_ * 4, 0);
        }
        const double unit = double(span_) * (8.0 / 30) * 1.3 * kWheelFrac;
        const double target = std::max((4.0 * nlarge_ + unit * inv_large) / (8.0 * E.cus), 2e4);
        std::vector<uint32_t> vc{0};
        for (uint32_t j = 0; j < nlarge_;) {
            const uint32_t hi = std::min<uint32_t>(nlarge_, j + GBH_HL_CHUNK);
            const double ij = E.inv_at(j);
            uint32_t e = std::min(j + 256, hi);
            while (e < hi) {   // grow by 256 primes while within the target
                const uint32_t nx = std::min(e + 256, hi);
                const double inx = nx == hi ? E.inv_at(nx) : E.inv_coarse[nx >> 8];
                if (4.0 * (nx - j) + unit * (inx - ij) > target && (nx & 255) == 0) break;
                e = nx;
            }
            vc.push_back(e);
            j = e;
        }
        nvchunks_ = uint32_t(vc.size() - 1);
        S.h_vchunks.ensure(c, vc.size() * 4, 1);
        std::memcpy(S.h_vchunks.p, vc.data(), vc.size() * 4);
        for (auto& B : S.sl) {
            B.vchunks.ensure(c, vc.size() * 4, 0);
            check(gbh_copy_h2d(c, 0, B.vchunks.p, S.h_vchunks.p, vc.size() * 4), "upload verify chunks");
            plan_copies_ = true;
        }
    }
    // the plan's uploads read pinned host tables that later plans rewrite
    if (plan_copies_) check(gbh_stream_sync(c, 0), "plan sync");
    plan_copies_ = false;
}

// Harvest chunks: contiguous prime ranges of about equal cost (a fixed cost
// per prime and run plus the expected hits; a wheel step costs about as
// much as a hit), enough chunks to fill every CU several times, 256 to
// GBH_HL_CHUNK primes each, split where the bitmap primes start and where a
// run's primes end. Uploaded for the batch in @p slot (stream-ordered after
// the batch before it).
void Pipeline::plan_chunks(const std::vector<RunCost>& runs_in, uint32_t nbb, double ncb, int slot)
{
    ChunkPlan& cp = E.cplan;
    if (cp.valid && cp.nbb == nbb && close_rel(cp.ncb, ncb) && cp.nwheel == nwheel_ && cp.nwheel_hi == nwheel_hi_ &&
        cp.nwheel_hi2 == nwheel_hi2_ && cp.nrand == nrand_ && cp.nx == nx_ && cp.mode == hl_mode_ &&
        cp.region_fix == region_fix_ && cp.chunks_cu == chunks_cu_ && same_shape(cp.key, runs_in)) {
        nchunks_ = cp.nchunks;
        nfew_ = cp.nfew;
        nsingle_ = cp.nsingle;
        stage_ = cp.stage;
        direct_ = cp.direct;
        ngroups_ = cp.ngroups;
        part_cap_ = cp.part_cap;
        upload_chunks(slot);   // unless the slot's tables hold it already
        return;
    }
    cp.valid = false;
    chunks_.assign(1, nx_);   // large primes below nx_ are sieved in the windows
    std::vector<double> hpb;    // per chunk: expected hits per class bit (the larger rate of its ends)
    std::vector<RunCost> rs;
    for (const RunCost& r : runs_in)
        if (r.np) rs.push_back(r);
    std::sort(rs.begin(), rs.end(), [](const RunCost& a, const RunCost& b) { return a.np < b.np; });
    stage_ = 0;
    if (!rs.empty()) {
        // Cost units per prime and per hit, by the way a prime is walked (a
        // prime of the single-multiple scan = 4; measured per-work-group
        // times on one MI355X, bench workload): bitmap primes with at most
        // four multipliers 7.5, other bitmap primes 8, those whose words are
        // a random read (likely a TLB miss across the bitmap) 32, wheel
        // primes 11 plus 2.5 per listed multiple; every hit 1 more.
        constexpr double kPrimeCost = 4.0, kFewCost = 7.5, kBitmapCost = 8.0, kRandCost = 32.0;
        constexpr double kWheelCost = 11.0, kWheelHitCost = 2.5;
        const size_t nr = rs.size();
        const uint32_t npmax = rs.back().np;
        {   // the compacted path's classes, for the largest run's q span (3.75 q-units per class bit)
            double span_q = 0;
            for (const RunCost& r : rs) span_q = std::max(span_q, 3.75 * r.bits);
            const uint32_t* lp = E.base.data() + E.win_hi;
            const auto first_above = [&](double x) {
                return uint32_t(std::upper_bound(lp, lp + npmax, uint32_t(std::min(x, 4294967295.0))) - lp);
            };
            nfew_ = std::min(ps_up(first_above(span_q / 4)), npmax);
            nsingle_ = std::min(ps_up(first_above(span_q)), npmax);
        }
        const auto prime_cost = [&](uint32_t j) {
            return on_wheel(j) ? kWheelCost : random_reads(j) ? kRandCost : j >= nsingle_ ? kPrimeCost
                 : j >= nfew_ ? kFewCost : kBitmapCost;
        };
        const auto hit_cost = [&](uint32_t j) { return on_wheel(j) ? 1.0 + kWheelHitCost : 1.0; };
        // runs [k, nr) are the ones serving primes below rs[k].np: their
        // summed class bits, largest bits and smallest first q
        std::vector<double> suf_bits(nr + 1, 0.0), suf_max(nr + 1, 0.0), suf_qlo(nr + 1, 1e30);
        for (size_t k = nr; k-- > 0;) {
            suf_bits[k] = suf_bits[k + 1] + rs[k].bits;
            suf_max[k] = std::max(suf_max[k + 1], rs[k].bits);
            suf_qlo[k] = std::min(suf_qlo[k + 1], rs[k].q_lo);
        }
        // sum of 1/p up to j: coarse (every 256 primes) where chunk ends fall, exact elsewhere
        const auto inv = [&](uint32_t j) {
            return (j & 255) == 0 ? E.inv_coarse[j >> 8]
                 : j == nwheel_ ? inv_wheel_ : j == nwheel_hi_ ? inv_hi_ : j == nwheel_hi2_ ? inv_hi2_ : E.inv_at(j);
        };
        const auto factor = [&](uint32_t a, double q_lo) { return on_wheel(a) ? 1.3 * kWheelFrac : hit_factor(a, q_lo); };
        double total = 0;
        {
            size_t ai = 0;
            for (uint32_t k = 0; k < 256; ++k) {
                uint32_t a = uint32_t(uint64_t(npmax) * k / 256) & ~255u, b = uint32_t(uint64_t(npmax) * (k + 1) / 256) & ~255u;
                if (k == 255) b = npmax;
                a = std::max(a, nx_);
                while (a < b) {
                    while (rs[ai].np <= a) ++ai;
                    const uint32_t e = next_cut(a, std::min(b, rs[ai].np));
                    const double q = suf_qlo[ai];
                    total += prime_cost(a) * (e - a) * double(nr - ai) +
                             hit_cost(a) * suf_bits[ai] * std::max(factor(a, q), factor(e - 1, q)) * (inv(e) - inv(a));
                    a = e;
                }
            }
        }
        // several chunks per CU (four at a time resident): the last ones to
        // start are short, so the kernel ends close to its total work / slots
        const double per_cu = chunks_cu_ ? chunks_cu_ : nbb <= 32 ? 16.0 : 8.0;
        const double target = std::max(total / (per_cu * E.cus), 2e4);
        // short batches' hits fit the work-groups' LDS stages: chunks of at
        // most kStageHits expected hits per run, one walk per prime
        constexpr double kStageHits = 0.7 * GBH_HL_STAGE;
        // long batches' direct regions: a chunk's largest (segment, class)
        // list well inside the small partition's LDS (0.9 of the lambda with
        // lambda + 8 sqrt(lambda) + 64 = GBH_PART_CAP_SMALL; chunks at their
        // 256-prime minimum may exceed it: then the large partition)
        const double x_cap = (-8.0 + std::sqrt(64.0 + 4.0 * (GBH_PART_CAP_SMALL - 64))) / 2;
        const double kRegionHits = 0.9 * x_cap * x_cap;
        double all_hits = 0;
        for (const RunCost& r : rs) all_hits += r.bits * r.density;
        stage_ = hl_mode_ != "direct" && all_hits <= kStageHits * 8192 ? GBH_HL_STAGE : 0u;
        uint32_t j = nx_, prev = 0;   // prev: the last chunk's length (lengths change slowly: start the search there)
        size_t ai = 0;
        while (j < npmax) {
            while (rs[ai].np <= j) ++ai;   // runs [ai, nr) serve prime j
            // (a run's last prime ends a chunk at the next group: the
            // kernel bounds every run's primes itself)
            const uint32_t hi = next_cut(j, std::min({npmax, j + GBH_HL_CHUNK, ps_up(rs[ai].np)}));
            const double nact = double(nr - ai), bits = suf_bits[ai], bmax = suf_max[ai], q = suf_qlo[ai];
            const double pc = prime_cost(j);
            const double hf = hit_factor(j, q);
            const double f = (on_wheel(j) ? 1.3 * kWheelFrac : hf) * hit_cost(j), fh = hf;
            const double ij = j == nwheel_ ? inv_wheel_ : E.inv_at(j);
            const auto ok = [&](uint32_t e) {   // chunk [j, e) within the targets (the rates barely change inside)
                const double si = inv(e) - ij;
                return pc * (e - j) * nact + bits * f * si <= target && (!stage_ || bmax * fh * si <= kStageHits) &&
                       (stage_ || ncb * fh * si <= kRegionHits);
            };
            uint32_t e;
            if (ok(hi)) {
                e = hi;
            } else {   // the last multiple of 256 that keeps the chunk within the targets (at least 256 primes)
                uint32_t lo = ((j >> 8) + 1) << 8, up = ((hi - 1) >> 8) << 8;
                if (lo > up) {
                    e = hi;
                } else {
                    if (lo < j + 256) lo += 256;
                    if (lo > up) lo = up;
                    // ok() grows with e: walk from the previous length to the
                    // largest e in [lo, up] with ok(e) (lo when none)
                    e = std::min(up, std::max(lo, ((j + prev) >> 8) << 8));
                    if (ok(e)) {
                        while (e < up && ok(e + 256)) e += 256;
                    } else {
                        while (e > lo && !ok(e)) e -= 256;
                    }
                }
            }
            prev = e - j;
            hpb.push_back(std::max(hf, hit_factor(e - 1, q)) * (inv(e) - ij));
            chunks_.push_back(e);
            j = e;
        }
    }
    nchunks_ = uint32_t(chunks_.size() - 1);
    // the direct harvest (batches whose hits exceed the LDS stages): work-group
    // k's region holds r_k entries per (segment, class) list, r_k well above
    // the largest list's expected hits lambda_k (lambda + 8 sqrt(lambda) + 64;
    // an overflow re-runs the batch with counted lists)
    direct_ = nchunks_ && !stage_ && hl_mode_ != "counted";
    std::vector<uint32_t>& rd = cp.rdesc;
    std::vector<uint32_t>& groups = cp.groups;   // the partition's chunk groups: their regions fit its LDS
    rd.clear();
    groups.clear();
    if (direct_) {
        rd.assign(2 * size_t(nchunks_), 0u);
        uint64_t base = 0, rmax = 0;
        for (uint32_t k = 0; k < nchunks_; ++k) {
            const double lam = ncb * hpb[k];
            const uint64_t r = region_fix_ ? region_fix_ : uint64_t(lam + 8 * std::sqrt(lam) + 64);
            rd[2 * k] = uint32_t(base);
            rd[2 * k + 1] = uint32_t(r);
            base += r * nbb;
            rmax = std::max(rmax, r);
            if (base >= UINT32_MAX || r > GBH_PART_CAP) {
                if (phase_trace_)
                    std::fprintf(stderr, "plan: direct harvest off (chunk %u of %u: region %llu entries per list, "
                                 "%llu in all)\n", k, nchunks_, (unsigned long long)r, (unsigned long long)base);
                direct_ = false;
                break;
            }
        }
        // the partition's groups: the small work-groups when every region fits
        part_cap_ = rmax <= GBH_PART_CAP_SMALL ? GBH_PART_CAP_SMALL : GBH_PART_CAP;
        groups.assign(1, 0u);
        uint64_t gsum = 0;
        for (uint32_t k = 0; direct_ && k < nchunks_; ++k) {
            const uint64_t r = rd[2 * k + 1];
            if (k > groups.back() && (k - groups.back() >= GBH_PART_CHUNKS || gsum + r > part_cap_)) {
                groups.push_back(k);
                gsum = 0;
            }
            gsum += r;
        }
        groups.push_back(nchunks_);
        if (direct_ && S.budget) {
            // a budgeted slot set: the regions (each slot's grown to the
            // larger of its present and its new size) must fit beside the
            // buffers held, else the counted form (no regions)
            uint64_t held = S.device_bytes(), after = held;
            for (const auto& B : S.sl) {
                const std::pair<const Buf*, uint64_t> grow[] = {
                    {&B.region, base * 4}, {&B.fills, uint64_t(nchunks_) * nbb * 4}, {&B.rdesc, rd.size() * 4},
                    {&B.groups, groups.size() * 4}, {&B.chunks, chunks_.size() * 4}};
                for (const auto& [buf, n] : grow) after += std::max<uint64_t>(n, 256) > buf->bytes ?
                                                               std::max<uint64_t>(n, 256) - buf->bytes : 0;
            }
            if (after > S.budget) {
                if (phase_trace_)
                    std::fprintf(stderr, "plan: direct harvest off (regions of %llu bytes beyond the device budget)\n",
                                 (unsigned long long)(after - held));
                direct_ = false;
            }
        }
        if (direct_) {
            region_entries_ = base;
            ngroups_ = uint32_t(groups.size() - 1);
            if (phase_trace_)
                std::fprintf(stderr, "plan: direct harvest: %llu region entries, at most %llu per list, %u partition "
                             "groups of up to %u entries\n", (unsigned long long)base, (unsigned long long)rmax,
                             ngroups_, part_cap_);
        } else {
            rd.clear();
            groups.clear();
        }
    }
    cp.chunks = chunks_;
    cp.region_entries = region_entries_;
    cp.key = runs_in;
    cp.nbb = nbb;
    cp.ncb = ncb;
    cp.nwheel = nwheel_;
    cp.nwheel_hi = nwheel_hi_;
    cp.nwheel_hi2 = nwheel_hi2_;
    cp.nrand = nrand_;
    cp.nx = nx_;
    cp.mode = hl_mode_;
    cp.region_fix = region_fix_;
    cp.chunks_cu = chunks_cu_;
    cp.nfew = nfew_;
    cp.nsingle = nsingle_;
    cp.nchunks = nchunks_;
    cp.stage = stage_;
    cp.direct = direct_;
    cp.ngroups = ngroups_;
    cp.part_cap = part_cap_;
    cp.valid = true;
    ++E.cplan_version;
    upload_chunks(slot);
}

// The current chunk plan into the device tables of @p slot (on the slot's
// stream, after the slot's batch before, which the host has collected).
void Pipeline::upload_chunks(int slot)
{
    if (S.slot_plan[slot] == E.cplan_version) return;
    const ChunkPlan& cp = E.cplan;
    gbh_ctx* c = E.c;
    auto& B = S.sl[slot];
    const size_t nc = cp.chunks.size(), nr = cp.rdesc.size(), ng = cp.groups.size();
    S.h_chunks[slot].ensure(c, (nc + nr + ng) * 4, 1);
    uint32_t* h = S.h_chunks[slot].as<uint32_t>();
    std::memcpy(h, cp.chunks.data(), nc * 4);
    B.chunks.ensure(c, nc * 4, 0);
    const int ms = main_stream(slot);
    check(gbh_copy_h2d(c, ms, B.chunks.p, h, nc * 4), "upload chunks");
    plan_copies_ = true;
    if (cp.direct) {
        std::memcpy(h + nc, cp.rdesc.data(), nr * 4);
        std::memcpy(h + nc + nr, cp.groups.data(), ng * 4);
        B.rdesc.ensure(c, nr * 4, 0);
        B.groups.ensure(c, ng * 4, 0);
        B.region.ensure(c, cp.region_entries * 4, 0);
        B.fills.ensure(c, size_t(cp.nchunks) * cp.nbb * 4, 0);
        check(gbh_copy_h2d(c, ms, B.rdesc.p, h + nc, nr * 4), "upload regions");
        check(gbh_copy_h2d(c, ms, B.groups.p, h + nc + nr, ng * 4), "upload region groups");
    }
    S.slot_plan[slot] = E.cplan_version;
}

// Segments of one range a batch may hold: the run's q range below 2^33 and
// its window lists within the harvest work-group's counters.
uint32_t Pipeline::run_cap(const Range& r) const
{
    const uint64_t kmax = std::max<uint64_t>(1, ((uint64_t(1) << 33) - 4 * MAX_GB_PRIME) / r.span - 1);
    return uint32_t(std::min<uint64_t>({kmax, GBH_HL_BUCKETS / (8 * nwin_max_), kb_}));
}

// The next batch: waiting segments of the ranges in submission order, each
// range's in order, up to kb_ segments (batched ranges: and about E.kbatch
// full segments of evens).
void Pipeline::form(std::vector<SegJob>& bt)
{
    bt.clear();
    uint64_t evens = 0;
    bool full = false;
    for (Range* r : active_) {
        uint32_t taken = 0;
        const uint32_t cap = run_cap(*r);
        while (!r->issued_all && bt.size() < kb_ && taken < cap) {
            const SegmentGeometry sg = segment_geometry(r->next_start, r->limit, r->span);
            if (multi_ && !bt.empty() && evens + sg.seg_evens > evens_cap_) { full = true; break; }
            SegJob j;
            j.r = r;
            j.sg = sg;
            j.index = r->next_index;
            j.sampled = r->stats_on && j.index % r->sample == 0;
            j.walked = r->gaps_on && (j.sampled || r->gaps_exact);
            j.counted = r->stats_on;
            bt.push_back(j);
            evens += sg.seg_evens;
            ++taken;
            ++r->next_index;
            r->rs.total_large_items += r->nlarge;
            if (!r->started) {
                r->started = true;
                r->t_first = clock_::now();
            }
            if (sg.seg_end == r->limit) r->issued_all = true;
            else r->next_start += r->span;
        }
        if (full || bt.size() >= kb_) break;
    }
    active_.erase(std::remove_if(active_.begin(), active_.end(), [](const Range* r) { return r->issued_all; }),
                  active_.end());
}

bool Pipeline::step(std::vector<Range*>* done)
{
    std::vector<SegJob>& bt = batches_[cur_];
    form(bt);
    if (!bt.empty()) enqueue(bt, cur_, true);
    if (pending_) {
        std::vector<SegJob>& prev = batches_[cur_ ^ 1];
        collect(prev, cur_ ^ 1);
        for (const SegJob& j : prev)
            if (j.r->last_collected && !j.r->finished) {
                finish_range(*j.r);
                if (done) done->push_back(j.r);
            }
    }
    pending_ = !bt.empty();
    cur_ ^= 1;
    return pending_ || !active_.empty();
}

void Pipeline::enqueue(std::vector<SegJob>& batch, int slot, bool check_and_count)
{
    gbh_ctx* c = E.c;
    auto& B = S.sl[slot];
    // streams: the slot's main stream (uploads, harvest, hit lists, checker,
    // counter) and its side stream (set-up and the window sieve phase)
    const int ms = main_stream(slot);
    const uint32_t nseg = uint32_t(batch.size());
    const TabLayout tl = tab_layout(nseg);
    uint8_t* ht = S.h_tab[slot].as<uint8_t>();
    gbh_geom* hg = reinterpret_cast<gbh_geom*>(ht);
    gbh_hl_run* hr = reinterpret_cast<gbh_hl_run*>(ht + tl.runs);
    gbh_countseg* hc = reinterpret_cast<gbh_countseg*>(ht + tl.cs);
    gbh_geom* dg = B.tab.as<gbh_geom>();
    gbh_hl_run* dr = reinterpret_cast<gbh_hl_run*>(B.tab.as<uint8_t>() + tl.runs);
    gbh_countseg* dcs = reinterpret_cast<gbh_countseg*>(B.tab.as<uint8_t>() + tl.cs);
    // batched ranges: a batch of short segments takes a window just above
    // its longest class array (one window per class); plain runs keep wbits_
    uint32_t wbits = wbits_;
    if (multi_ && !E.win_kb) {
        uint64_t ncm = 1;
        for (SegJob& j : batch) {
            j.cg = class_geometry(j.sg.q_first, j.sg.q_odds);
            for (int k = 0; k < 8; ++k) ncm = std::max<uint64_t>(ncm, j.cg.n[k]);
        }
        uint32_t w = 1u << 13;
        while (w < wbits_ && w < ncm) w <<= 1;
        wbits = w;
    }
    uint32_t nwin = 1, nwords = 0, nruns = 0, run_max = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
        SegJob& j = batch[s];
        Range& r = *j.r;
        j.cg = class_geometry(j.sg.q_first, j.sg.q_odds);
        gbh_geom g = make_geom(j.sg, j.cg, wbits);
        // evens below q_first + 9780 (only at the very bottom of the range)
        // have candidate q outside the map: the CPU resolves them
        uint64_t n_gpu = std::max<uint64_t>(j.sg.seg_start, j.sg.q_first + 9780);
        n_gpu += n_gpu & 1;
        g.n_gpu = n_gpu;
        j.n_gpu = n_gpu;
        g.hist_on = j.sampled ? 1u : 0u;
        g.np_win = r.nwin + std::min(nx_, r.nlarge);
        // ridge records: a campaign run records from its store's current
        // threshold; a run without campaign files from the floor (the same
        // records however its segments are batched)
        g.ridge_t = !r.ridges || !r.ridges->active() ? 0xFFFFFFFFu : r.campaign ? r.ridges->thresh() : r.ridges->floor();
        j.gpu_evens = n_gpu <= j.sg.seg_end ? (j.sg.seg_end - n_gpu) / 2 + 1 : 0;
        if (g.nwin > nwin_max_) throw std::runtime_error("segment longer than the run's buffers");
        nwin = std::max(nwin, g.nwin);
        hg[s] = g;
        gbh_countseg cs{};
        if (j.counted) {
            cs = make_count_args(j.sg, j.cg);
            nwords = std::max(nwords, cs.nwords);
        }
        hc[s] = cs;
        if (!force_count_) r.rs.total_tiles += uint64_t((j.cg.n[0] + wbits - 1) / wbits) * 8;   // once per batch
        if (s == 0 || batch[s - 1].r != j.r) {
            hr[nruns++] = gbh_hl_run{s, 0, r.nlarge, 0};
        }
        ++hr[nruns - 1].nseg;
        run_max = std::max(run_max, hr[nruns - 1].nseg);
    }
    gbh_batch b{dg, nseg, cw_, wbits, nwin, B.map.as<uint64_t>()};
    // this batch's output block and chunk counter; the other pair is cleared
    // for the slot's next batch beside this one's harvest
    const uint32_t sel = S.out_sel[slot];
    S.out_sel[slot] = sel ^ 1u;
    gbh_segout* out = S.d_out[slot][sel].as<gbh_segout>();
    uint32_t* hq = B.hq.as<uint32_t>() + 32 * sel;
    auto& ev = S.ev[slot];
    captured_[slot] = graph_;
    if (graph_) {
        check(gbh_event_record(c, ms, S.tbeg[slot].e), "event");
        check(gbh_graph_begin(c, ms), "graph capture");
    }
    mark(ms, slot, EV_BEGIN);
    harvested_[slot] = nlarge_ != 0;
    std::vector<RunCost> costs;
    {   // algorithmic bytes of this batch's harvest and window launches, per run
        double bits_all = 0, hits_all = 0, win_prime_bytes = 0;
        for (uint32_t k = 0; k < nruns; ++k) {
            const gbh_hl_run& run = hr[k];
            const Range& r = *batch[run.seg0].r;
            const double lq = double(hg[run.seg0 + run.nseg - 1].q_last - hg[run.seg0].q_first);
            double bits = 0;
            for (uint32_t s = run.seg0; s < run.seg0 + run.nseg; ++s)
                for (int x = 0; x < 8; ++x) bits += hg[s].nc[x];
            const double hits = r.density * bits;
            // the run's bitmap primes: [nwheel_, nwheel_hi_) and [nwheel_hi2_, np).
            // Words a prime must read: below the run's q span one per prime
            // plus its multiplier range (q span / p / 240); above it, one
            // only when its single multiple lies in the span and its
            // multiplier is a unit mod 30 ((8/30) q span / p); so too from a
            // quarter of the span on when the compacted path takes those
            // primes (one word per multiple whose multiplier is a unit)
            const uint32_t np = r.nlarge, a0 = std::min(np, nwheel_), a1 = std::min(np, nwheel_hi_),
                           b0 = std::min(np, nwheel_hi2_);
            const uint32_t* lp = E.base.data() + E.win_hi;
            const auto first_above = [&](double v) {
                return uint32_t(std::upper_bound(lp, lp + np, uint32_t(std::min(v, 4294967295.0))) - lp);
            };
            const uint32_t js = first_above(lq);
            const uint32_t jf = few_path() && direct_ && !force_count_ ? first_above(lq / 4) : js;
            double words = 0;
            for (const auto& [x, y] : {std::pair<uint32_t, uint32_t>{a0, a1}, {b0, np}}) {
                if (x >= y) continue;
                const uint32_t f = std::clamp(std::min(jf, js), x, y);
                const double ix = E.inv_at(x), i_f = E.inv_at(f), iy = E.inv_at(y);
                words += double(f - x) + lq * (i_f - ix) / 240 + (8.0 / 30) * lq * (iy - i_f);
            }
            // the packed primes: one byte each and a 4-byte base per group
            if (!force_count_ && np > nx_) {
                agg_.harvest_bytes += (1.0 + 4.0 / GBH_PS_GROUP) * (np - nx_) + 8.0 * words + 4.0 * hits;
                agg_.harvest_words += words;
            }
            bits_all += bits;
            hits_all += hits;
            win_prime_bytes += 4.0 * run.nseg * 8 * (r.nwin + std::min(nx_, r.nlarge));
            costs.push_back(RunCost{np, bits, double(r.q_lo), r.density});
        }
        if (!force_count_) agg_.window_bytes += bits_all / 8 + win_prime_bytes + 4.0 * hits_all;
    }
    if (multi_ && nlarge_) {
        double ncb = 0;
        for (uint32_t s = 0; s < nseg; ++s)
            for (int x = 0; x < 8; ++x) ncb = std::max<double>(ncb, hg[s].nc[x]);
        plan_chunks(costs, nseg * 8, ncb, slot);
    } else if (nlarge_) {
        upload_chunks(slot);
    }
    // the harvest first (the longest stage): the tables, the error word and
    // the fill counts it needs, then its launch; the sieve phase and the
    // rest of the set-up are queued behind it
    check(gbh_copy_h2d(c, ms, B.tab.p, ht, tl.bytes), "upload batch tables");
    const bool zeroed = S.clean_n[slot][sel] >= nseg && S.clean_gen[slot][sel][0] == S.d_out[slot][sel].gen &&
                        S.clean_gen[slot][sel][1] == B.hq.gen;
    S.clean_n[slot][sel] = 0;
    if (!zeroed) check(gbh_reset_outputs(c, ms, out, nseg), "reset outputs");
    uint32_t ncoop = 0;
    while (ncoop < nwin_ && E.base[E.win_lo + ncoop] < wbits / 128) ++ncoop;
    // the window stage's sieve phase (masks, window primes) on the slot's
    // side stream beside the harvest, once the batch tables are uploaded
    const int ss = overlap_ ? 2 + slot : ms;
    mark(ms, slot, EV_GEOM);
    const bool direct = nlarge_ && direct_ && !force_count_;
    direct_used_[slot] = direct;
    if (nlarge_) {
        // the counted and staged forms reserve list slots in hit_n; the
        // direct form needs it only for the partition (zeroed on the side
        // stream, below) and writes every region fill count itself
        if (!direct) check(gbh_zero(c, ms, B.hit_n.p, size_t(nseg) * 8 * nwin * 4), "zero hit counts");
        gbh_harvest_args h{b, E.large, B.chunks.as<uint32_t>(), dr, nruns,
                           run_max, nchunks_, cap_, t_hl_, t_hi_, t_hi2_, direct ? 0u : stage_,
                           mb_ ? mb_->bits : nullptr,
                           mb_ ? mb_->m_base : 0, mb_ ? mb_->m_span : 0, B.hits.as<uint32_t>(),
                           B.hit_n.as<uint32_t>(), direct ? B.rdesc.as<uint32_t>() : nullptr,
                           direct ? B.region.as<uint32_t>() : nullptr, direct ? B.fills.as<uint32_t>() : nullptr,
                           &out[0].err, persist_ ? hq : nullptr, persist_wg_,
                           few_path() ? 1u : 0u, zeroed ? 1u : 0u, hit_queue_ && few_path() ? 1u : 0u};
        mark(ms, slot, EV_HL0);
        check(gbh_launch_harvest(c, ms, &h), "harvest");
        mark(ms, slot, EV_HL1);
    }
    if (ss != ms) check(gbh_stream_wait(c, ss, ev[EV_GEOM].e), "stream wait");
    // the partition's list counts, zeroed beside the harvest (the main
    // stream waits for EV_S0 before the partition)
    if (direct) check(gbh_zero(c, ss, B.hit_n.p, size_t(nseg) * 8 * nwin * 4), "zero hit counts");
    // the checker's candidate tables depend on the geometry alone: built here
    // on the side stream, ahead of the set-up (the main stream's wait for the
    // set-up before the partition covers them)
    const uint32_t cgroups = nseg < 4 ? 3u : 1u;   // short batches: the 15 even classes over 3 work-groups
    gbh_check_args ca{b, E.ka, pcap_, B.coff.as<uint32_t>(), B.coffa.as<uint32_t>(), B.pending.as<uint32_t>(), out,
                      0u, cgroups};
    if (check_and_count) check(gbh_launch_check_a(c, ss, &ca), "check candidate tables");
    const uint32_t nwp = nwin_ + nx_;   // the windows' primes
    if (nwp) {
        gbh_setup_args a{b, E.d_win.as<uint32_t>(), nwp, 0, B.i0.as<uint32_t>()};
        check(gbh_launch_setup(c, ss, &a), "setup");
    }
    mark(ss, slot, EV_S0);
    {
        const gbh_window_args w{b, E.d_win.as<uint32_t>(), nwp, ncoop, B.i0.as<uint32_t>(), nullptr, nullptr, 0, 1};
        check(gbh_launch_window(c, ss, &w), "window sieve");
    }
    {   // the slot's next batch: its output block and chunk counter, cleared
        // here beside the harvest (EV_S1, which the hit-list phase waits for,
        // covers them, so they are done before this batch ends)
        const uint32_t nx = sel ^ 1u;
        check(gbh_reset_outputs(c, ss, S.d_out[slot][nx].as<gbh_segout>(), kb_), "reset outputs");
        check(gbh_fill32(c, ss, B.hq.as<uint32_t>() + 32 * nx, 1, 0u), "clear the harvest chunk counter");
        S.clean_n[slot][nx] = kb_;
        S.clean_gen[slot][nx][0] = S.d_out[slot][nx].gen;
        S.clean_gen[slot][nx][1] = B.hq.gen;
    }
    mark(ss, slot, EV_S1);
    if (direct) {
        if (ss != ms) check(gbh_stream_wait(c, ms, ev[EV_S0].e), "stream wait");
        const gbh_partition_args pa{b, B.rdesc.as<uint32_t>(), B.region.as<uint32_t>(), B.fills.as<uint32_t>(),
                                    B.groups.as<uint32_t>(), ngroups_, cap_, B.hits.as<uint32_t>(),
                                    B.hit_n.as<uint32_t>(), &out[0].err, part_cap_, 0};
        check(gbh_launch_partition(c, ms, &pa), "partition");
    }
    mark(ms, slot, EV_SETUP);
    if (ss != ms) check(gbh_stream_wait(c, ms, ev[EV_S1].e), "stream wait");
    if (nlarge_) {   // the hit-list phase: every window's list applied to the sieved map
        const gbh_window_args w{b, E.d_win.as<uint32_t>(), nwp, ncoop, B.i0.as<uint32_t>(),
                                B.hits.as<uint32_t>(), B.hit_n.as<uint32_t>(), cap_, 2};
        check(gbh_launch_window(c, ms, &w), "window hit lists");
    }
    mark(ms, slot, EV_WINDOW);
    count_side_[slot] = ss != ms && check_and_count && nwords;
    if (count_side_[slot]) check(gbh_stream_wait(c, ss, ev[EV_WINDOW].e), "stream wait");
    if (verify_every_ && nlarge_ && check_and_count)
        for (uint32_t s = 0; s < nseg; ++s)
            if (batch[s].index % verify_every_ == 0) enqueue_verify(s, b, out, slot);
    // both kinds with the counter beside the checker (split): phase A of
    // the sampled segments and then phase B on the side stream, phase A of
    // the others and then the counter here (the counter's VALU work beside
    // phase B's latency-bound probes); else phase B after both phase A's
    // here and the counter on the side stream from the window's end
    bool split = false;
    if (check_and_count) {
        uint32_t kinds = 0;
        for (const SegJob& j : batch) kinds |= j.sampled ? 2u : 1u;
        ca.kinds = kinds | 4;                // the candidate tables are built
        gbh_checkb_args cb{b, E.ka, CHECK_TAIL, B.coff.as<uint32_t>(), B.pending.as<uint32_t>(), pcap_, kend_, out,
                           B.ridge_more.as<uint64_t>()};
        if (kinds == 3 && count_side_[slot]) {
            // both kinds: the histogram segments on the side stream, the
            // others here, phase B after both
            split = checkb_split_ && nwords;
            ca.kinds = 1 | 4;
            check(gbh_launch_check_a(c, ms, &ca), "check phase A");
            if (split) mark(ms, slot, EV_CA);
            ca.kinds = 2 | 4;
            if (cgroups == 3) ca.class_groups = 5;   // the slower histogram form: more waves for it
            check(gbh_launch_check_a(c, ss, &ca), "check phase A (histogram)");
            if (split) {
                check(gbh_stream_wait(c, ss, ev[EV_CA].e), "stream wait");
                check(gbh_launch_check_b(c, ss, &cb), "check phase B");
                mark(ss, slot, EV_CHECK);
            } else {
                mark(ss, slot, EV_CA);
                check(gbh_stream_wait(c, ms, ev[EV_CA].e), "stream wait");
            }
        } else {
            check(gbh_launch_check_a(c, ms, &ca), "check phase A");
        }
        if (!split) check(gbh_launch_check_b(c, ms, &cb), "check phase B");
    }
    if (!split) mark(ms, slot, EV_CHECK);
    if (check_and_count && nwords) {
        // the counter only reads the map and fills its own output fields:
        // on the side stream it runs beside the checker
        const int cs = count_side_[slot] && !split ? ss : ms;
        gbh_count_args ct{b, dcs, nwords, swords_, B.scratch.as<uint32_t>(), out};
        check(gbh_launch_count(c, cs, &ct), "count");
        mark(cs, slot, EV_COUNT);
        if (cs != ms) check(gbh_stream_wait(c, ms, ev[EV_COUNT].e), "stream wait");
    } else {
        mark(ms, slot, EV_COUNT);
    }
    if (split) check(gbh_stream_wait(c, ms, ev[EV_CHECK].e), "stream wait");
    // the blocks without their leftover lists (fetched in collect when used)
    check(gbh_copy_d2h_rows(c, ms, S.h_out[slot].p, sizeof(gbh_segout), out, sizeof(gbh_segout), GBH_SEGOUT_HEAD,
                            nseg), "copy outputs");
    if (graph_) {   // the captured batch: one launch
        check(gbh_graph_end(c, ms, S.graph[slot].g), "graph launch");
        ++agg_.graph_launches;
    }
    check(gbh_event_record(c, ms, ev[EV_END].e), "event");
    if (phase_trace_) std::fprintf(stderr, "host: batch queued %.3f ms after its first segment was formed\n",
                                   since(batch[0].r->t_first) * 1e3);
}

// The GB_BUCKET_VERIFY rebuild of segment s (the reference's
// enqueue_bucket_verify, src/verifier.cpp:2059-2085): window primes as in
// production, every large prime by the wheel walk (no bitmap), then a word
// compare with the production map into out[s].vdiff.
void Pipeline::enqueue_verify(uint32_t s, const gbh_batch& b, gbh_segout* out, int slot)
{
    gbh_ctx* c = E.c;
    auto& B = S.sl[slot];
    const int ms = main_stream(slot);
    const gbh_batch vb{b.geoms + s, 1, cw_, b.wbits, b.nwin, B.vmap.as<uint64_t>()};
    uint32_t ncoop = 0;
    while (ncoop < nwin_ && E.base[E.win_lo + ncoop] < b.wbits / 128) ++ncoop;
    if (nwin_) {
        const gbh_setup_args a{vb, E.d_win.as<uint32_t>(), nwin_, 0, B.i0.as<uint32_t>()};
        check(gbh_launch_setup(c, ms, &a), "verify setup");
    }
    check(gbh_zero(c, ms, B.vhit_n.p, size_t(8) * b.nwin * 4), "verify zero");
    const gbh_harvest_args h{vb, E.large, B.vchunks.as<uint32_t>(), nullptr, 0, 0, nvchunks_, cap_v_,
                             UINT32_MAX, UINT32_MAX, UINT32_MAX, 0, nullptr, 0, 0, B.vhits.as<uint32_t>(), B.vhit_n.as<uint32_t>(),
                             nullptr, nullptr, nullptr, &out[s].err, nullptr, 0, 0, 0, 0};
    check(gbh_launch_harvest(c, ms, &h), "verify harvest");
    const gbh_window_args w{vb, E.d_win.as<uint32_t>(), nwin_, ncoop, B.i0.as<uint32_t>(),
                            B.vhits.as<uint32_t>(), B.vhit_n.as<uint32_t>(), cap_v_, 0};
    check(gbh_launch_window(c, ms, &w), "verify window");
    const gbh_mapcmp_args m{B.map.as<uint64_t>() + size_t(s) * 8 * cw_, B.vmap.as<uint64_t>(),
                            b.geoms + s, cw_, 0, &out[s].vdiff};
    check(gbh_launch_mapcmp(c, ms, &m), "verify compare");
    ++verified_;
}

void Pipeline::collect(std::vector<SegJob>& batch, int slot)
{
    gbh_ctx* c = E.c;
    auto& ev = S.ev[slot];
    const auto tw = clock_::now();
    check(gbh_event_sync(c, ev[EV_END].e), "batch sync");
    const auto tcol = clock_::now();
    if (phase_trace_) std::fprintf(stderr, "host: waited %.3f ms for the batch\n", since(tw) * 1e3);
    gbh_segout* outs = S.h_out[slot].as<gbh_segout>();
    {   // the leftover lists of the segments that have any (the CPU fallback)
        const gbh_segout* dev = S.d_out[slot][S.out_sel[slot] ^ 1u].as<gbh_segout>();   // this batch's block
        bool any = false;
        for (size_t s = 0; s < batch.size(); ++s) {
            const uint32_t n = std::min<uint32_t>(outs[s].left_n, GBH_LEFT_CAP);
            if (!n) continue;
            check(gbh_copy_d2h(c, main_stream(slot), outs[s].left, dev[s].left, size_t(n) * 4), "copy leftovers");
            any = true;
        }
        if (any) check(gbh_stream_sync(c, main_stream(slot)), "leftover sync");
    }
    if ((outs[0].err & 1u) && direct_used_[slot]) {
        // a region of the direct harvest overflowed: the batch again, with
        // counted lists (queued behind the batch in flight, which keeps its
        // own outputs); the failed attempt's times and bytes are not counted
        std::fprintf(stderr, "note: harvest region overflow in the batch from %llu: re-run with counted lists\n",
                     (unsigned long long)batch[0].sg.seg_start);
        ++agg_.harvest_reruns;
        force_count_ = true;
        enqueue(batch, slot, true);
        force_count_ = false;
        collect(batch, slot);
        return;
    }
    // the sieve phase ran beside the harvest: its own time plus the hit-list
    // phase count as the window stage
    // stage times from the events between the stages; a captured batch has
    // only its whole time (launch to end), counted as its harvest stage
    const auto t = [&](int id) -> const Event& { return ev[id]; };
    const bool cap = captured_[slot];
    const double tl = cap ? elapsed_s(c, S.tbeg[slot], ev[EV_END]) : elapsed_s(c, t(EV_BEGIN), t(EV_SETUP));
    const double ts = cap ? 0.0 : elapsed_s(c, t(EV_S0), t(EV_S1)) + elapsed_s(c, t(EV_SETUP), t(EV_WINDOW));
    if (harvested_[slot] && !cap) {
        agg_.harvest_kernel_seconds += elapsed_s(c, t(EV_HL0), t(EV_HL1));
        ++agg_.harvest_launches;
    }
    if (!cap) {
        agg_.window_kernel_seconds += ts;
        ++agg_.window_launches;
    }
    const double tc = cap ? 0.0 : elapsed_s(c, t(EV_WINDOW), t(EV_CHECK));
    // the counter ran beside the checker (from the window's end) or after it
    const double tn = cap ? 0.0 : elapsed_s(c, t(count_side_[slot] ? EV_WINDOW : EV_CHECK), t(EV_COUNT));
    agg_.sieve_seconds += ts;
    agg_.large_seconds += tl;
    agg_.verify_seconds += tc + tn;
    agg_.count_seconds += tn;
    if (outs[0].err & 1u) throw std::runtime_error("hit-list bucket overflow (raise the capacity)");
    if (outs[0].err & 4u) throw std::runtime_error("hit lists: a bitmap prime's cube is not above the batch's q range");
    if (outs[0].err & 8u) throw std::runtime_error("hit lists: a multiplier outside the multiplier bitmap");
    if (outs[0].err & 16u) throw std::runtime_error("hit lists: a quotient estimate more than one off");
    const double share[4] = {ts / batch.size(), tl / batch.size(), tc / batch.size(), tn / batch.size()};
    for (size_t s = 0; s < batch.size(); ++s) collect_segment(batch[s], outs[s], share, slot, uint32_t(s));
    if (phase_trace_) std::fprintf(stderr, "host: collected in %.3f ms\n", since(tcol) * 1e3);
    if (phase_trace_)
        std::fprintf(stderr, "batch of %zu segments from %llu: harvest+setup %.3f ms window %.3f ms check %.3f ms "
                     "count %.3f ms\n", batch.size(), (unsigned long long)batch[0].sg.seg_start, tl * 1e3, ts * 1e3,
                     tc * 1e3, tn * 1e3);
}

void Pipeline::collect_segment(const SegJob& j, const gbh_segout& o, const double share[4], int slot, uint32_t s)
{
    Range& r = *j.r;
    r.t_sieve += share[0];
    r.t_large += share[1];
    r.t_check += share[2];
    r.t_count += share[3];
    const std::string seg = " in segment [" + std::to_string(j.sg.seg_start) + ", " + std::to_string(j.sg.seg_end) + "]";
    if (o.err & 1u) throw std::runtime_error("hit-list bucket overflow (raise the capacity)" + seg);
    if (o.err & ~13u) throw std::runtime_error("device error flags " + std::to_string(o.err) + seg);
    if (o.vdiff)
        throw std::runtime_error("GB_BUCKET_VERIFY: " + std::to_string(o.vdiff) + " map words differ from the "
                                 "wheel-only rebuild" + seg);
    if (o.pending_n > pcap_) throw std::runtime_error("pending overflow (" + std::to_string(o.pending_n) + ")" + seg);
    // coverage: every even accounted for, by phase A, phase B or a leftover
    if (uint64_t(o.acct[0]) + o.pending_n != j.gpu_evens)
        throw std::runtime_error("coverage: phase A resolved " + std::to_string(o.acct[0]) + " + pending " +
                                 std::to_string(o.pending_n) + " != " + std::to_string(j.gpu_evens) + " evens" + seg);
    if (uint64_t(o.acct[1]) + o.left_n != o.pending_n)
        throw std::runtime_error("coverage: phase B resolved " + std::to_string(o.acct[1]) + " + leftovers " +
                                 std::to_string(o.left_n) + " != pending " + std::to_string(o.pending_n) + seg);
    if (o.left_n > GBH_LEFT_CAP || o.left_n > MAX_LEFTOVERS) throw std::runtime_error("leftover overflow" + seg);
    // the CPU fallback: leftovers, and the evens below the GPU's first checkable even
    std::vector<uint64_t> cpu_n;
    for (uint64_t n = j.sg.seg_start; n <= j.sg.seg_end && n < j.n_gpu; n += 2) cpu_n.push_back(n);
    for (uint32_t k = 0; k < o.left_n; ++k) {
        const uint32_t i = o.left[k];
        if (i >= j.sg.seg_evens) throw std::runtime_error("leftover index " + std::to_string(i) + " outside the segment");
        cpu_n.push_back(j.sg.seg_start + 2ull * i);
    }
    std::sort(cpu_n.begin(), cpu_n.end());
    for (uint64_t n : cpu_n) {
        const uint64_t p = cpu_min_goldbach_prime(n, E.cand.primes.data(), E.cand.primes.size(), MAX_GB_PRIME);
        if (!p) throw std::runtime_error("Goldbach verification FAILED for n=" + std::to_string(n));
        if (p > MAX_GB_PRIME)
            std::fprintf(stderr, "note: n=%llu needs p=%llu > MAX_GB_PRIME (new p_min record candidate)\n",
                         (unsigned long long)n, (unsigned long long)p);
        if (p > 2 && r.gs) r.gs->add_cpu(n, p, j.sampled);
        if (p > 2 && r.ridges) r.ridges->consider(n, p, j.sg.seg_start);
    }
    r.rs.leftovers += o.left_n;
    if (r.ridges) {
        // the first GBH_RIDGE_CAP records came with the output block; more
        // (a low GB_RIDGE_MIN) wait in the slot's overflow buffer
        const uint32_t kept = std::min<uint32_t>(o.ridge_n, GBH_RIDGE_CAP + GBH_RIDGE_MORE);
        if (kept > GBH_RIDGE_CAP) {
            std::vector<uint64_t> all(o.ridge, o.ridge + 2 * GBH_RIDGE_CAP);
            all.resize(2 * size_t(kept));
            const uint64_t* src = S.sl[slot].ridge_more.as<uint64_t>() + 2 * size_t(s) * GBH_RIDGE_MORE;
            const int ms = main_stream(slot);
            check(gbh_copy_d2h(E.c, ms, all.data() + 2 * GBH_RIDGE_CAP, src, size_t(kept - GBH_RIDGE_CAP) * 16),
                  "copy ridge records");
            check(gbh_stream_sync(E.c, ms), "ridge records sync");
            r.ridges->merge(all.data(), kept, o.ridge_n, j.sg.seg_start);
        } else {
            r.ridges->merge(o.ridge, kept, o.ridge_n, j.sg.seg_start);
        }
    }
    if (r.gs) {
        r.gs->add_hist_slots(j.sampled, o.hist, GBH_NCAND, E.cand.slot_pi);
        r.gs->add_first(j.sg.seg_start, o.first);
        if (r.gaps_on) {
            if (o.gtail_n > GBH_GAP_TCAP) throw std::runtime_error("count: gap tail list overflow" + seg);
            SegmentGaps g;
            g.q_first = j.sg.q_first;
            g.evens = j.sg.seg_evens;
            for (int k = 0; k < 8; ++k) g.primes += o.cnt[k];
            g.pfirst = o.pfirst;
            g.plast = o.plast;
            g.tail = o.gtail;
            g.ntail = o.gtail_n;
            g.walked = j.walked;
            g.gmax = o.gmax;
            g.hist = o.ghist;
            r.gs->add_gap_segment(g);
        }
        r.gs->add_counts(o.cnt, o.twins);
        if (j.sampled) r.sampled_total += j.sg.seg_evens;
        r.evens_done += j.sg.seg_evens;
        r.gs->write_live(r.start, r.limit, r.evens_done, r.total_evens, since(r.t_submit), false);
    }
    if (r.verbose)
        std::printf("  seg %llu: [%llu, %llu] evens=%llu n_left=%u\n", (unsigned long long)j.index,
                    (unsigned long long)j.sg.seg_start, (unsigned long long)j.sg.seg_end,
                    (unsigned long long)j.sg.seg_evens, o.left_n);
    if (r.collect) {
        SegmentStats st;
        st.seg_start = j.sg.seg_start;
        st.seg_end = j.sg.seg_end;
        st.evens = j.sg.seg_evens;
        st.tiles = uint64_t((j.cg.n[0] + wbits_ - 1) / wbits_) * 8;
        st.large_items = r.nlarge;
        st.sieve_split = r.split;
        st.mask_count = E.win_lo;
        st.leftovers = o.left_n;
        st.sieve_seconds = share[0];
        st.large_seconds = share[1];
        st.check_seconds = share[2] + share[3];
        r.rs.per_segment.push_back(st);
    }
    if (j.sg.seg_end == r.limit) r.last_collected = true;
}

// A range whose last segment is collected: timings, ridges, the statistics
// commit (campaign files) or the in-memory state (batched and served ranges).
void Pipeline::finish_range(Range& r)
{
    RunStats& rs = r.rs;
    rs.segments = r.next_index;
    rs.gpu_seconds = since(r.t_first);
    rs.init_seconds = std::chrono::duration<double>(r.t_first - r.t_submit).count();
    rs.total_seconds = rs.init_seconds + rs.gpu_seconds;
    rs.evens_per_second = double(r.total_evens) / rs.gpu_seconds;
    rs.sieve_seconds = r.t_sieve;
    rs.large_seconds = r.t_large;
    rs.verify_seconds = r.t_check + r.t_count;
    rs.count_seconds = r.t_count;
    rs.sync_seconds = std::max(0.0, rs.gpu_seconds - r.t_sieve - r.t_large - r.t_check - r.t_count);
    rs.start = r.start;
    if (!multi_) {   // a plain run owns every launch
        rs.harvest_kernel_seconds = agg_.harvest_kernel_seconds;
        rs.window_kernel_seconds = agg_.window_kernel_seconds;
        rs.harvest_launches = agg_.harvest_launches;
        rs.harvest_reruns = agg_.harvest_reruns;
        rs.graph_launches = agg_.graph_launches;
        rs.harvest_words = agg_.harvest_words;
        rs.window_launches = agg_.window_launches;
        rs.harvest_bytes = agg_.harvest_bytes;
        rs.window_bytes = agg_.window_bytes;
        if (verify_every_ && nlarge_)
            std::fprintf(stderr, "bucket verify: %llu segment maps rebuilt without the multiplier bitmap, identical\n",
                         (unsigned long long)verified_);
    }
    if (r.ridges && r.ridges->active()) {
        if (r.campaign) r.ridges->save(r.ridge_path);
        rs.ridge_min = r.ridges->floor();
        rs.ridge_found = r.ridges->appends();
        rs.ridge_dropped = r.ridges->dropped();
        rs.ridge_top = r.ridges->top();
        if (r.ridges->overflowed())
            std::fprintf(stderr, "warning: a segment recorded more than %u ridges at or above the threshold; the "
                         "excess was not considered\n", GBH_RIDGE_CAP + GBH_RIDGE_MORE);
    }
    if (r.gs) {
        r.gs->finish_run(r.start, r.limit, r.total_evens);
        rs.stats_on = true;
        rs.stats_pmax = r.gs->run_pmax();
        rs.stats_primes = r.gs->run_primes();
        rs.stats_twins = r.gs->run_twins();
        rs.stats_sampled = r.gs->evens_sampled();
        rs.stats_tail = r.gs->run_tail();
        r.gs->write_live(r.start, r.limit, r.total_evens, r.total_evens, since(r.t_submit), true);
        rs.stats_line = r.gs->log_line(VERSION);
        const auto tc = clock_::now();
        if (r.campaign) r.gs->commit(VERSION, r.sampled_total, &rs.stats_line);
        else r.state_text = r.gs->memory_commit(r.sampled_total);
        if (phase_trace_) std::fprintf(stderr, "host: statistics commit %.3f ms\n", since(tc) * 1e3);
    }
    r.finished = true;
}

uint32_t Pipeline::map_of(Range& r, uint64_t seg_start, std::vector<uint64_t>& words, uint64_t Q[8], uint32_t nc[8])
{
    std::vector<SegJob> bt(1);
    bt[0].r = &r;
    bt[0].sg = segment_geometry(seg_start, r.limit, r.span);
    enqueue(bt, 0, false);
    check(gbh_event_sync(E.c, S.ev[0][EV_END].e), "sync");
    if (S.h_out[0].as<gbh_segout>()[0].err & 13u) throw std::runtime_error("hit-list device error");
    words.assign(size_t(8) * cw_, 0);
    check(gbh_copy_d2h(E.c, 0, words.data(), S.sl[0].map.p, words.size() * 8), "copy map");
    check(gbh_stream_sync(E.c, 0), "sync");
    for (int c = 0; c < 8; ++c) { Q[c] = bt[0].cg.Q[c]; nc[c] = uint32_t(bt[0].cg.n[c]); }
    return cw_;
}

// One plain run over [start, limit] (the reference's run_gpu body).
RunStats run_plain(Engine::State& E, uint64_t limit, uint64_t segment_even, bool verbose, uint64_t start_even,
                   uint32_t split_arg, bool collect)
{
    const auto t0 = clock_::now();
    const bool trace = env::present("GB_PHASE");
    Range r;
    r.limit = limit;
    r.span = segment_even;
    r.start = start_even;
    r.verbose = verbose;
    r.collect = collect;
    r.t_submit = t0;
    resolve_range(E, r, split_arg);
    Pipeline P(E, false);
    const double t_res = since(t0);
    configure_stats(E, r);
    configure_ridges(E, r);
    const double t_cfg = since(t0);
    PlanBounds pb = bounds_of(r);
    P.plan(pb);
    r.density = P.density_of(r.q_lo, r.nlarge);
    r.rs.host_plan_seconds = since(t0);
    if (trace)
        std::fprintf(stderr, "host: resolve %.3f ms, statistics set-up %.3f ms, plan %.3f ms\n", t_res * 1e3,
                     (t_cfg - t_res) * 1e3, (r.rs.host_plan_seconds - t_cfg) * 1e3);
    P.add(&r);
    while (P.step(nullptr)) {}
    if (!r.finished) throw std::runtime_error("run ended before its last segment was collected");
    return std::move(r.rs);
}

}  // namespace

// ---------------------------------------------------------------- batches of ranges

struct Batcher::Impl {
    Engine::State& E;
    Slots slots;                       // this batcher's batch buffers (within its device budget)
    Pipeline P;
    uint64_t max_limit, span, min_start;
    std::vector<std::unique_ptr<Range>> ranges;
    std::vector<Result> results;
    Impl(Engine::State& e, uint64_t max_limit_, uint64_t min_start_, uint64_t segment_even, uint64_t budget)
        : E(e), P(e, true, &slots), max_limit(max_limit_ & ~uint64_t(1)), span(segment_even + (segment_even & 1))
    {
        slots.init(E.c, 1);
        slots.budget = budget;
        if (max_limit > E.max_limit) throw std::runtime_error("batcher limit above the engine's prime bound");
        if (span < 4 || span > E.max_span) throw std::runtime_error("batcher segment span outside the engine's buffers");
        min_start = std::max<uint64_t>(min_start_ + (min_start_ & 1), 4);
        if (min_start > max_limit) throw std::runtime_error("batcher: no even in [min_start, max_limit]");
        // the bounds every range must fit: primes of max_limit, the bitmap
        // from the smallest start up, lists sized for the densest magnitude
        Range top;
        top.limit = max_limit;
        top.start = min_start;
        top.span = span;
        resolve_range(E, top, 0);
        PlanBounds pb = bounds_of(top);
        pb.span = span;
        pb.nseg = GBH_MAX_BATCH;
        pb.q_lo_np.clear();
        const double l0 = std::log(double(std::max<uint64_t>(min_start, 16))), l1 = std::log(double(max_limit));
        for (int k = 0; k <= 24; ++k) {
            const double x = std::exp(l0 + (l1 - l0) * k / 24);
            const uint64_t q = uint64_t(std::min(x, double(max_limit)));
            const uint32_t bound = std::max<uint32_t>(base_prime_bound(std::max<uint64_t>(q + 2 * span, 64)), 64);
            const uint32_t nb = uint32_t(std::upper_bound(E.base.begin(), E.base.end(), bound) - E.base.begin());
            pb.q_lo_np.push_back({q > MAX_GB_PRIME ? q - MAX_GB_PRIME : 3, nb > E.win_hi ? nb - E.win_hi : 0});
        }
        P.plan(pb);
    }
    ~Impl()
    {
        // nothing of this batcher's may still run when its buffers go
        for (int s = 0; s < 4; ++s) (void)gbh_stream_sync(E.c, s);
    }
    // the segments [start_even, limit] splits into (as a plain run splits it)
    uint64_t segments_of(uint64_t start_even, uint64_t limit) const
    {
        Range r;
        r.limit = limit;
        r.start = start_even;
        r.span = span;
        resolve_range(E, r, 0);
        uint64_t n = 0;
        for (uint64_t s = r.start;; s += r.span) {
            ++n;
            if (segment_geometry(s, r.limit, r.span).seg_end == r.limit) break;
        }
        return n;
    }
};

Batcher::Batcher(Engine& e, uint64_t max_limit, uint64_t min_start, uint64_t segment_even, uint64_t device_budget)
    : p_(std::make_unique<Impl>(*e.state(), max_limit, min_start, segment_even, device_budget))
{
}
Batcher::~Batcher() = default;

size_t Batcher::submit(uint64_t start_even, uint64_t limit)
{
    Impl& I = *p_;
    auto r = std::make_unique<Range>();
    r->id = I.ranges.size();
    r->limit = limit;
    r->start = start_even;
    r->span = I.span;
    r->campaign = false;
    r->t_submit = clock_::now();
    if ((limit & ~uint64_t(1)) > I.max_limit) throw std::runtime_error("range limit above the batcher's bound");
    if (start_even < I.min_start) throw std::runtime_error("range start below the batcher's bound");
    resolve_range(I.E, *r, 0);
    configure_stats(I.E, *r);
    configure_ridges(I.E, *r);
    r->density = I.P.density_of(r->q_lo, r->nlarge);
    I.P.add(r.get());
    I.ranges.push_back(std::move(r));
    I.results.emplace_back();
    return I.ranges.size() - 1;
}

bool Batcher::step(std::vector<size_t>* finished)
{
    Impl& I = *p_;
    std::vector<Range*> done;
    const bool more = I.P.step(&done);
    for (Range* r : done) {
        Result& res = I.results[r->id];
        res.stats = std::move(r->rs);
        res.state = std::move(r->state_text);
        res.done = true;
        r->gs.reset();
        r->ridges.reset();
        if (finished) finished->push_back(r->id);
    }
    return more;
}

const Batcher::Result& Batcher::result(size_t id) const
{
    if (id >= p_->results.size()) throw std::runtime_error("no such range");
    return p_->results[id];
}

size_t Batcher::size() const { return p_->ranges.size(); }
uint32_t Batcher::batch_segments() const { return p_->P.kb(); }
uint64_t Batcher::device_bytes() const { return p_->slots.device_bytes(); }
uint64_t Batcher::segments_of(uint64_t start_even, uint64_t limit) const { return p_->segments_of(start_even, limit); }
const RunStats& Batcher::totals() const { return p_->P.totals(); }

Engine::Engine(uint64_t max_limit, uint64_t max_span, const char* prime_cache)
    : s_(std::make_unique<State>(max_limit, max_span, prime_cache))
{
}
Engine::~Engine() = default;
double Engine::init_seconds() const { return s_->init_s; }
double Engine::prepare(uint64_t limit, uint64_t start_even)
{
    const auto t = clock_::now();
    Range r;
    r.limit = limit;
    r.start = start_even;
    r.span = s_->max_span;
    resolve_range(*s_, r, 0);
    Pipeline P(*s_, false);
    P.plan_bitmap(bounds_of(r));
    return since(t);
}
uint64_t Engine::max_limit() const { return s_->max_limit; }
void Engine::init_times(double& cache, double& module, double& upload) const
{
    cache = s_->cache_s;
    module = s_->module_s;
    upload = s_->upload_s;
}

RunStats Engine::run(uint64_t limit, uint64_t segment_even, bool verbose, uint64_t start_even, uint32_t sieve_split_arg,
                     bool collect_segments)
{
    return run_plain(*s_, limit, segment_even, verbose, start_even, sieve_split_arg, collect_segments);
}

namespace {
// The span a run over [start, limit] keeps after the range resolution.
uint64_t resolved_span(uint64_t limit, uint64_t start_even, uint64_t segment_even)
{
    const uint64_t lim = limit & ~uint64_t(1);
    const uint64_t st = std::max<uint64_t>(start_even + (start_even & 1), 4);
    uint64_t span = std::max<uint64_t>(segment_even + (segment_even & 1), 2);
    if (st <= lim) span = std::min<uint64_t>(span, std::max<uint64_t>((((lim - st) / 2 + 1) & ~uint64_t(1)), 4));
    return std::max<uint64_t>(span, 4);
}
}  // namespace

RunStats run_gpu(uint64_t limit, uint64_t segment_even, const std::string& spv_path, bool verbose, uint64_t start_even,
                 uint32_t sieve_split_arg, const char* prime_cache_path, bool collect_segments)
{
    (void)spv_path;
    refuse_removed_knobs();
    if (limit < 4) throw std::runtime_error("limit < 4: no even n >= 4 to verify");
    require_supported_limit(limit);
    Engine e(limit & ~uint64_t(1), resolved_span(limit, start_even, segment_even), prime_cache_path);
    RunStats s = e.run(limit, segment_even, verbose, start_even, sieve_split_arg, collect_segments);
    s.init_seconds += e.init_seconds();
    s.total_seconds = s.init_seconds + s.gpu_seconds;
    e.init_times(s.init_cache_s, s.init_module_s, s.init_upload_s);
    return s;
}

void print_device_info()
{
    Ctx c;
    gbh_device_info_t i{};
    check(gbh_device_info(c.c, &i), "gbh_device_info");
    std::printf("device 0: %s (%s)\n  compute units : %d\n  LDS per block : %zu bytes\n  HBM           : %.1f GB\n"
                "  clock         : %d MHz\n  window primes < %llu, phase A %u candidates\n",
                i.name, i.arch, i.cus, i.lds_per_block, double(i.hbm_bytes) / 1e9, i.clock_khz / 1000,
                (unsigned long long)WIN_P_DEFAULT, KA_DEFAULT);
}

bool verify_cpu(uint64_t limit, bool verbose)
{
    const auto prime = [](uint64_t n) {
        if (n < 2) return false;
        if (n < 4) return true;
        if (n % 2 == 0) return false;
        for (uint64_t d = 3; d <= n / d; d += 2)
            if (n % d == 0) return false;
        return true;
    };
    for (uint64_t n = 4; n <= limit; n += 2) {
        bool ok = false;
        for (uint64_t p = 2; p <= n / 2 && !ok; ++p) ok = prime(p) && prime(n - p);
        if (!ok) {
            if (verbose) std::printf("FAIL at n=%llu\n", (unsigned long long)n);
            return false;
        }
        if (verbose && n % 1000000 == 0) std::printf("  %llu ok\n", (unsigned long long)n);
    }
    return true;
}

uint32_t gpu_segment_map(uint64_t seg_start, uint64_t limit, uint64_t span, std::vector<uint64_t>& words,
                         uint64_t Q[8], uint32_t nc[8])
{
    if (seg_start & 1 || limit & 1 || seg_start < 4 || seg_start > limit) throw std::runtime_error("bad segment");
    Engine e(limit, std::max<uint64_t>(span, 4));
    Range r;
    r.limit = limit;
    r.span = span;
    r.start = seg_start;
    resolve_range(*e.state(), r, 0);
    Pipeline P(*e.state(), false);
    P.plan(bounds_of(r));
    r.density = P.density_of(r.q_lo, r.nlarge);
    return P.map_of(r, seg_start, words, Q, nc);
}

}  // namespace gb
