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

enum { EV_BEGIN, EV_HL0, EV_HL1, EV_SETUP, EV_WINDOW, EV_CHECK, EV_COUNT, EV_END, EV_N };

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

// One segment of a batch, host side.
struct SegJob {
    SegmentGeometry sg;
    ClassGeo cg;
    uint64_t index = 0;
    uint64_t n_gpu = 0;
    uint64_t gpu_evens = 0;
    bool sampled = false, walked = false, counted = false;
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
    Buf d_win, d_large, d_chunks, d_i0, d_map, d_coff, d_coffa, d_pending, d_scratch, d_hits, d_hit_n;
    Buf d_vmap, d_vhits, d_vhit_n, d_vchunks, h_vchunks;   // the GB_BUCKET_VERIFY rebuild
    Buf d_geoms, d_cs;
    Buf d_out[2], h_out[2], h_geoms[2], h_cs[2], h_chunks;
    Event ev[2][EV_N];
    Candidates cand;

    State(uint64_t max_limit_, uint64_t max_span_, const char* prime_cache)
    {
        t0 = clock_::now();
        require_supported_limit(max_limit_);
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
        const uint64_t win_p = env::u64("GB_WIN_P", WIN_P_DEFAULT, 67, UINT32_MAX);
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
        d_win.ensure(c, size_t(nwin) * 4, 0);
        d_large.ensure(c, size_t(nlarge) * 4, 0);
        if (nwin) check(gbh_copy_h2d(c, 0, d_win.p, base.data() + win_lo, size_t(nwin) * 4), "upload window primes");
        if (nlarge) check(gbh_copy_h2d(c, 0, d_large.p, base.data() + win_hi, size_t(nlarge) * 4), "upload large primes");
        for (int s = 0; s < 2; ++s) {
            d_out[s].ensure(c, sizeof(gbh_segout) * kbatch, 0);
            h_out[s].ensure(c, sizeof(gbh_segout) * kbatch, 1);
            h_geoms[s].ensure(c, sizeof(gbh_geom) * kbatch, 1);
            h_cs[s].ensure(c, sizeof(gbh_countseg) * kbatch, 1);
            for (auto& e : ev[s]) e.create(c);
        }
        d_geoms.ensure(c, sizeof(gbh_geom) * kbatch, 0);
        d_cs.ensure(c, sizeof(gbh_countseg) * kbatch, 0);
        d_coff.ensure(c, size_t(kbatch) * 15 * GBH_NCAND * 4, 0);
        d_coffa.ensure(c, size_t(kbatch) * 15 * GBH_NCAND_PAD * 4, 0);
        check(gbh_stream_sync(c, 0), "upload sync");
        upload_s = since(tu);
        init_s = since(t0);
    }
};

namespace {

// One run over [start, limit] on an engine's resident state (the
// reference's DenseRun phases: resolve_range, statistics set-up, the
// segment loop, finish and commit), issued in batches of segments.
class DenseRun {
public:
    DenseRun(Engine::State& e, uint64_t limit, uint64_t segment_even, bool verbose, uint64_t start_even,
             uint32_t split_arg, bool collect)
        : E(e), limit_(limit), span_(segment_even), verbose_(verbose), start_(start_even), split_arg_(split_arg),
          collect_(collect)
    {
    }
    RunStats run();
    /// Range resolution and the multiplier bitmap only; the build's seconds.
    double prepare()
    {
        const auto t = clock_::now();
        resolve_range();
        return since(t);
    }
    /// The class map of segment [seg_start, ...] alone (validators).
    uint32_t map_of(uint64_t seg_start, std::vector<uint64_t>& words, uint64_t Q[8], uint32_t nc[8]);

private:
    void resolve_range();
    void plan_bitmap();
    void configure_stats();
    void configure_ridges();
    void plan_buffers();
    double hit_factor(uint32_t j) const;
    double hit_density() const;
    double compute_hit_density() const;
    double hit_density_ = 0;
    void enqueue(std::vector<SegJob>& batch, int slot, bool check_and_count);
    void collect(std::vector<SegJob>& batch, int slot);
    void collect_segment(const SegJob& j, const gbh_segout& o, double share[4]);

    Engine::State& E;
    uint64_t limit_, span_;
    bool verbose_;
    uint64_t start_;
    uint32_t split_arg_;
    bool collect_;
    RunStats rs_;
    uint64_t total_evens_ = 0;
    clock_::time_point t_start_;
    uint32_t nwin_ = 0, nlarge_ = 0, split_ = 0, ncoop_ = 0;   // this run's window / large prime counts
    uint32_t wbits_ = 0, nwin_max_ = 0, cw_ = 0, cap_ = 0, pcap_ = 0, swords_ = 0, nchunks_ = 0, kb_ = 1;
    // large primes: [0, nwheel_) walk the wheel, [nwheel_, nlarge_) read the multiplier bitmap from t_hl_ up
    uint32_t nwheel_ = 0, t_hl_ = UINT32_MAX;
    uint32_t stage_ = 0;       // harvest LDS stage per work-group (short batches: one walk per prime)
    // GB_BUCKET_VERIFY=N: every Nth segment's map is rebuilt with every large
    // prime on the wheel walk (no multiplier bitmap) and compared word for word
    uint32_t verify_every_ = 0, cap_v_ = 0, nvchunks_ = 0;
    uint64_t verified_ = 0;
    void enqueue_verify(uint32_t s, const gbh_batch& b, gbh_segout* out);
    double inv_wheel_ = 0, inv_large_ = 0;   // sum of 1/p over the wheel primes, over all large primes
    std::shared_ptr<const MultBitmap> mb_;
    bool stats_on_ = false, gaps_on_ = false, gaps_exact_ = false;
    uint32_t sample_ = 256;
    std::optional<GoldbachStats> gs_;
    std::optional<RidgeTracker> ridges_;
    std::string ridge_path_;
    uint64_t sampled_total_ = 0, evens_done_ = 0;
    bool phase_trace_ = false;
    bool harvested_[2] = {false, false};
    double t_sieve_ = 0, t_large_ = 0, t_check_ = 0, t_count_ = 0;
};

void DenseRun::resolve_range()
{
    refuse_removed_knobs();
    require_supported_limit(limit_);
    limit_ &= ~uint64_t(1);
    rs_.limit = limit_;
    if (limit_ < 4) throw std::runtime_error("limit < 4: no even n >= 4 to verify");
    if (limit_ > E.max_limit) throw std::runtime_error("limit above the engine's prime bound");
    if (start_ < 4) start_ = 4;
    if (start_ & 1) start_ += 1;
    if (start_ > limit_) throw std::runtime_error("no even n in [start, limit]");
    if (span_ < 2) span_ = 2;
    if (span_ & 1) span_ += 1;
    total_evens_ = (limit_ - start_) / 2 + 1;
    // at least two segments (the span is an integer span: span / 2 evens)
    if (span_ > total_evens_) span_ = std::max<uint64_t>(total_evens_ & ~uint64_t(1), 2);
    if (start_ == 4) {
        if (limit_ < 6)
            throw std::runtime_error("the GPU path needs limit >= 6 (n = 4 = 2 + 2 has no odd-prime partition; "
                                     "--cpu-check 4 verifies it)");
        span_ = std::max<uint64_t>(span_, 4);
    }
    if (span_ > E.max_span) throw std::runtime_error("segment span above the engine's buffers");
    // a batch's positions are 32-bit half-offsets: one merged segment (below
    // two spans) must stay below 2^33 integers
    if (span_ >= (uint64_t(1) << 32)) throw std::runtime_error("segment span of 2^32 integers or more");
    // this run's primes: those <= sqrt(limit) + 16
    const uint32_t bound = std::max<uint32_t>(base_prime_bound(limit_), 64);
    const uint32_t nb = uint32_t(std::upper_bound(E.base.begin(), E.base.end(), bound) - E.base.begin());
    uint32_t hi = std::min(E.win_hi, nb);
    if (split_arg_ > 0) {          // --split: the first large prime's index among the base primes
        if (split_arg_ < E.win_lo || split_arg_ > E.win_hi)
            throw std::runtime_error("--split must lie in [" + std::to_string(E.win_lo) + ", " +
                                     std::to_string(E.win_hi) + "] on this engine");
        if (split_arg_ < hi)
            std::fprintf(stderr, "note: --split below the engine's window bound: the windows keep their primes\n");
    }
    nwin_ = hi > E.win_lo ? hi - E.win_lo : 0;
    split_ = E.win_lo + nwin_;
    nlarge_ = nb > E.win_hi ? nb - E.win_hi : 0;
    rs_.split_used = split_;
    plan_bitmap();
}

// The large primes' split between the wheel walk and the multiplier bitmap:
// a prime p >= t_hl lists p m with m prime, exact while every q < t_hl^3;
// the bitmap covers m in [max(t_hl, sqrt(q_lo)), q_hi / t_hl] within its
// budget (t_hl rises until it fits).
void DenseRun::plan_bitmap()
{
    nwheel_ = nlarge_;
    t_hl_ = UINT32_MAX;
    mb_.reset();
    if (!nlarge_) return;
    const uint64_t budget = mult_bitmap_budget();
    if (!budget) return;
    const uint64_t q_lo = start_ >= MAX_GB_PRIME + 2 ? start_ - MAX_GB_PRIME : 3, q_hi = limit_ - 3;
    const uint64_t m_cap = (budget - std::min<uint64_t>(budget, uint64_t(1) << 24)) * 30;   // bitmap bytes -> m span
    const uint64_t sq_lo = uint64_t(std::sqrt(double(q_lo)));
    const uint32_t* lp = E.base.data() + E.win_hi;
    // smallest bitmap prime: its cube above q_hi, and the m span within the budget
    const auto cube_le = [](uint64_t x, uint64_t q) { return (unsigned __int128)x * x * x <= q; };
    uint64_t cb = uint64_t(std::cbrt(double(q_hi)));
    while (cube_le(cb + 1, q_hi)) ++cb;
    while (cb > 1 && !cube_le(cb, q_hi)) --cb;
    const uint64_t t = std::max<uint64_t>({lp[0], cb + 1, q_hi / (m_cap + sq_lo) + 1});
    nwheel_ = uint32_t(std::lower_bound(lp, lp + nlarge_, uint32_t(std::min<uint64_t>(t, UINT32_MAX))) - lp);
    if (nwheel_ == nlarge_) return;
    t_hl_ = lp[nwheel_];
    const uint64_t m_lo = std::max<uint64_t>(t_hl_, sq_lo > 1 ? sq_lo - 1 : 1), m_hi = q_hi / t_hl_;
    if (m_hi < m_lo) { nwheel_ = nlarge_; t_hl_ = UINT32_MAX; return; }
    mb_ = mult_bitmap(m_lo, m_hi, E.base, verbose_ || phase_trace_);
}

// Goldbach ridges (the reference's configure_ridges_and_stats,
// src/verifier.cpp:718-744): the floor is clamped above every phase-A
// candidate, so above it phase B sees every even; the store is loaded after
// the clamp.
void DenseRun::configure_ridges()
{
    ridges_.emplace(&E.cand.primes);
    ridge_path_ = env::path("GB_RIDGE_FILE", "goldbach_ridges.csv");
    uint32_t pa_max = 0;
    for (uint32_t cr = 0; cr < 15; ++cr)
        pa_max = std::max(pa_max, E.cand.primes[E.cand.slot_pi[cr][std::min<size_t>(E.ka, E.cand.slot_pi[cr].size()) - 1]]);
    if (ridges_->active()) {
        ridges_->clamp_floor_above(pa_max);
        ridges_->load(ridge_path_);
    }
}

void DenseRun::configure_stats()
{
    const bool stats_env = env::onoff("GB_STATS", true);
    stats_on_ = start_ >= STATS_MIN_START && stats_env;
    const std::string gaps_env = env::choice("GB_GAPS", "exact", {"exact", "1", "sampled", "0"});
    gaps_on_ = stats_on_ && gaps_env != "0";
    gaps_exact_ = gaps_on_ && gaps_env != "sampled";
    sample_ = env::u32("GB_STATS_SAMPLE", 256, 1);
    if (stats_on_) {
        gs_.emplace(E.cand.primes, CHECK_TAIL);
        gs_->set_run_start(start_);
        if (gaps_on_) gs_->begin_gaps(start_);
    } else if (start_ < STATS_MIN_START && !env::present("GB_STATS")) {
        std::fprintf(stderr, "note: Goldbach statistics off below start %llu\n", (unsigned long long)STATS_MIN_START);
    }
    // every phase-A candidate must lie below the tail bound: evens with
    // p(n) >= tail_t then all reach phase B, which records them exactly
    for (uint32_t cr = 0; cr < 15; ++cr)
        if (E.cand.primes[E.cand.slot_pi[cr][std::min<size_t>(E.ka, E.cand.slot_pi[cr].size()) - 1]] >= CHECK_TAIL)
            throw std::runtime_error("GB_KA too large: a phase-A candidate reaches the tail bound");
}

// Expected hits per class bit and unit 1/p of large prime j (the j-th above
// the window bound): the wheel lists every multiplier coprime to 7 11 13,
// the bitmap lists prime multipliers, (30/8) / ln m of those coprime to 30
// (m = q / p >= max(q_lo / p, p), the bound that makes the rate largest).
double DenseRun::hit_factor(uint32_t j) const
{
    if (j < nwheel_) return kWheelFrac;
    const double p = E.base[E.win_hi + j];
    const double q_lo = double(start_ >= MAX_GB_PRIME + 2 ? start_ - MAX_GB_PRIME : 3);
    return 3.75 / std::log(std::max({q_lo / p, p, 3.0}));
}

// Expected hits per class bit over all large primes: the wheel primes
// exactly, the bitmap primes in 64 bands of the prime index, each at its
// largest rate.
double DenseRun::hit_density() const { return hit_density_; }

double DenseRun::compute_hit_density() const
{
    if (!nlarge_) return 0.0;
    double d = kWheelFrac * E.inv_at(nwheel_);
    const uint32_t n = nlarge_ - nwheel_;
    for (uint32_t k = 0; k < 64 && n; ++k) {
        const uint32_t a = nwheel_ + uint32_t(uint64_t(n) * k / 64), b = nwheel_ + uint32_t(uint64_t(n) * (k + 1) / 64);
        if (b <= a) continue;
        d += (E.inv_at(b) - E.inv_at(a)) * std::max(hit_factor(a), hit_factor(b - 1));
    }
    return d;
}

// Window size, strides, hit-list capacity and buffer sizes of this run.
void DenseRun::plan_buffers()
{
    const uint64_t nseg = std::min<uint64_t>(E.kbatch, (total_evens_ + span_ / 2 - 1) / std::max<uint64_t>(span_ / 2, 1) + 1);
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
        while (wbits_ > (1u << 15) && nseg * 8 * ((nc_typ + wbits_ - 1) / wbits_) < 2ull * E.cus) wbits_ >>= 1;
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
    hit_density_ = compute_hit_density();
    const double lam = wbits_ * hit_density();
    cap_ = uint32_t(std::min(lam * 1.15 + 8 * std::sqrt(lam) + 1024, 4e9 / (double(E.kbatch) * 8 * nwin_max_)));
    const uint64_t evens_max = span_ + 2;
    pcap_ = uint32_t(std::min<uint64_t>(evens_max / 32 + 65536, UINT32_MAX / 2 / E.kbatch));
    swords_ = uint32_t(2 * (nc_max / 64 + 4));
    gbh_ctx* c = E.c;
    // a batch's q range stays below 2^33 (the harvest's 32-bit half-offsets):
    // at most kb_ - 1 spans plus the merged last segment (below 2 spans)
    const uint64_t kmax = std::max<uint64_t>(1, ((uint64_t(1) << 33) - 4 * MAX_GB_PRIME) / span_ - 1);
    kb_ = uint32_t(std::min<uint64_t>({nseg, kmax, GBH_HL_BUCKETS / (8 * nwin_max_)}));
    if (!kb_) throw std::runtime_error("segment span too large for the hit-list work-group");
    const size_t K = kb_;
    E.d_map.ensure(c, K * 8 * cw_ * 8, 0);
    E.d_i0.ensure(c, K * 8 * size_t(std::max(nwin_, 1u)) * 4, 0);
    E.d_pending.ensure(c, K * size_t(pcap_) * 4, 0);
    E.d_scratch.ensure(c, K * size_t(swords_) * 4, 0);
    E.d_hit_n.ensure(c, K * 8 * size_t(nwin_max_) * 4, 0);
    if (nlarge_) E.d_hits.ensure(c, K * 8 * size_t(nwin_max_) * cap_ * 4, 0);
    const auto tp0 = clock_::now();
    // harvest chunks: contiguous prime ranges of about equal cost (a fixed
    // cost per prime plus its expected hits; a wheel step costs about as
    // much as a hit), enough chunks to fill every CU several times, 256 to
    // GBH_HL_CHUNK primes each, split where the bitmap primes start
    std::vector<uint32_t> ch{0};
    inv_wheel_ = E.inv_at(nwheel_);
    inv_large_ = E.inv_at(nlarge_);
    if (nlarge_) {
        const double hits_unit = double(K) * double(span_) * (8.0 / 30);   // expected hits of prime p: unit * factor / p
        constexpr double kPrimeCost = 4.0;
        // sum of 1/p up to j: exact at the split and the end, coarse (every
        // 256 primes) elsewhere, where chunk ends fall
        const auto inv = [&](uint32_t j) {
            return (j & 255) == 0 ? E.inv_coarse[j >> 8]
                 : j == nwheel_ ? inv_wheel_ : j == nlarge_ ? inv_large_ : E.inv_at(j);
        };
        const auto factor = [&](uint32_t a) { return a < nwheel_ ? 1.3 * kWheelFrac : hit_factor(a); };
        double total = 0;
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t a = uint32_t(uint64_t(nlarge_) * k / 256) & ~255u, b = uint32_t(uint64_t(nlarge_) * (k + 1) / 256) & ~255u;
            if (k == 255) b = nlarge_;
            if (a < nwheel_ && b > nwheel_) {
                total += kPrimeCost * (nwheel_ - a) + hits_unit * factor(a) * (inv(nwheel_) - inv(a));
                a = nwheel_;
            }
            if (b > a) total += kPrimeCost * (b - a) + hits_unit * std::max(factor(a), factor(b - 1)) * (inv(b) - inv(a));
        }
        const double target = std::max(total / (8.0 * E.cus), 2e4);
        // a short batch's hits fit the work-groups' LDS stages: chunks of at
        // most kStageHits expected hits, one walk per prime
        constexpr double kStageHits = 0.7 * GBH_HL_STAGE;
        const double all_hits = hits_unit * hit_density();
        stage_ = all_hits <= kStageHits * 8192 ? GBH_HL_STAGE : 0u;
        uint32_t j = 0, prev = 0;   // prev: the last chunk's length (lengths change slowly: start the search there)
        while (j < nlarge_) {
            uint32_t hi = std::min<uint32_t>(nlarge_, j + GBH_HL_CHUNK);
            if (j < nwheel_) hi = std::min(hi, nwheel_);
            const double hf = hit_factor(j);
            const double f = j < nwheel_ ? 1.3 * kWheelFrac : hf, fh = hf;
            const double ij = j == nwheel_ ? inv_wheel_ : E.inv_at(j);
            const auto ok = [&](uint32_t e) {   // chunk [j, e) within the targets (the rates barely change inside)
                const double si = inv(e) - ij;
                return kPrimeCost * (e - j) + hits_unit * f * si <= target && (!stage_ || hits_unit * fh * si <= kStageHits);
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
            ch.push_back(e);
            j = e;
        }
    }
    const double t_chunks = since(tp0);
    nchunks_ = uint32_t(ch.size() - 1);
    if (phase_trace_)
        std::fprintf(stderr, "plan: %u segments per batch, window %u bits, %u window primes, %u large (%u on the wheel, "
                     "bitmap from %u), %u harvest chunks, list cap %u, stage %u\n", kb_, wbits_, nwin_, nlarge_,
                     nwheel_, t_hl_ == UINT32_MAX ? 0u : t_hl_, nchunks_, cap_, stage_);
    if (phase_trace_) std::fprintf(stderr, "plan: chunking %.3f ms\n", t_chunks * 1e3);
    E.h_chunks.ensure(c, ch.size() * 4, 1);
    E.d_chunks.ensure(c, ch.size() * 4, 0);
    std::memcpy(E.h_chunks.p, ch.data(), ch.size() * 4);
    check(gbh_copy_h2d(c, 0, E.d_chunks.p, E.h_chunks.p, ch.size() * 4), "upload chunks");
    verify_every_ = env::present("GB_BUCKET_VERIFY") ? env::u32("GB_BUCKET_VERIFY", 0, 0) : 0u;
    if (verify_every_ && nlarge_) {
        // one segment, every large prime on the wheel: its own lists and chunks
        const double lam_v = wbits_ * kWheelFrac * inv_large_;
        cap_v_ = uint32_t(lam_v * 1.15 + 8 * std::sqrt(lam_v) + 1024);
        E.d_vmap.ensure(c, size_t(8) * cw_ * 8, 0);
        E.d_vhits.ensure(c, size_t(8) * nwin_max_ * cap_v_ * 4, 0);
        E.d_vhit_n.ensure(c, size_t(8) * nwin_max_ * 4, 0);
        const double unit = double(span_) * (8.0 / 30) * 1.3 * kWheelFrac;
        const double target = std::max((4.0 * nlarge_ + unit * inv_large_) / (8.0 * E.cus), 2e4);
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
        E.h_vchunks.ensure(c, vc.size() * 4, 1);
        E.d_vchunks.ensure(c, vc.size() * 4, 0);
        std::memcpy(E.h_vchunks.p, vc.data(), vc.size() * 4);
        check(gbh_copy_h2d(c, 0, E.d_vchunks.p, E.h_vchunks.p, vc.size() * 4), "upload verify chunks");
    }
    check(gbh_stream_sync(c, 0), "chunk sync");
}

void DenseRun::enqueue(std::vector<SegJob>& batch, int slot, bool check_and_count)
{
    gbh_ctx* c = E.c;
    const uint32_t nseg = uint32_t(batch.size());
    gbh_geom* hg = E.h_geoms[slot].as<gbh_geom>();
    gbh_countseg* hc = E.h_cs[slot].as<gbh_countseg>();
    uint32_t nwin = 1, nwords = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
        SegJob& j = batch[s];
        j.cg = class_geometry(j.sg.q_first, j.sg.q_odds);
        gbh_geom g = make_geom(j.sg, j.cg, wbits_);
        // evens below q_first + 9780 (only at the very bottom of the range)
        // have candidate q outside the map: the CPU resolves them
        uint64_t n_gpu = std::max<uint64_t>(j.sg.seg_start, j.sg.q_first + 9780);
        n_gpu += n_gpu & 1;
        g.n_gpu = n_gpu;
        j.n_gpu = n_gpu;
        g.hist_on = j.sampled ? 1u : 0u;
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
    }
    gbh_batch b{E.d_geoms.as<gbh_geom>(), nseg, cw_, wbits_, nwin, E.d_map.as<uint64_t>()};
    gbh_segout* out = E.d_out[slot].as<gbh_segout>();
    auto& ev = E.ev[slot];
    check(gbh_event_record(c, 0, ev[EV_BEGIN].e), "event");
    harvested_[slot] = nlarge_ != 0;
    {   // algorithmic bytes of this batch's harvest and window launches
        const double lq = double(hg[nseg - 1].q_last - hg[0].q_first);
        double bits = 0;
        for (uint32_t s = 0; s < nseg; ++s)
            for (int k = 0; k < 8; ++k) bits += hg[s].nc[k];
        const double hits = hit_density() * bits;
        const double s_bm = nlarge_ ? inv_large_ - inv_wheel_ : 0.0;
        rs_.harvest_bytes += nlarge_ ? 4.0 * nlarge_ + 8.0 * ((nlarge_ - nwheel_) + lq * s_bm / 240) + 4.0 * hits : 0.0;
        rs_.window_bytes += bits / 8 + 4.0 * nseg * 8 * nwin_ + 4.0 * hits;
    }
    check(gbh_copy_h2d(c, 0, E.d_geoms.p, hg, sizeof(gbh_geom) * nseg), "upload geometry");
    check(gbh_copy_h2d(c, 0, E.d_cs.p, hc, sizeof(gbh_countseg) * nseg), "upload count parameters");
    check(gbh_reset_outputs(c, 0, out, nseg), "reset outputs");
    if (nwin_) {
        gbh_setup_args a{b, E.d_win.as<uint32_t>(), nwin_, 0, E.d_i0.as<uint32_t>()};
        check(gbh_launch_setup(c, 0, &a), "setup");
    }
    if (nlarge_) {
        check(gbh_zero(c, 0, E.d_hit_n.p, size_t(nseg) * 8 * nwin * 4), "zero hit counts");
        gbh_harvest_args h{b, E.d_large.as<uint32_t>(), E.d_chunks.as<uint32_t>(), nchunks_, cap_, t_hl_, stage_,
                           mb_ ? mb_->bits : nullptr, mb_ ? mb_->m_base : 0, mb_ ? mb_->m_span : 0,
                           E.d_hits.as<uint32_t>(), E.d_hit_n.as<uint32_t>(), &out[0].err};
        check(gbh_event_record(c, 0, ev[EV_HL0].e), "event");
        check(gbh_launch_harvest(c, 0, &h), "harvest");
        check(gbh_event_record(c, 0, ev[EV_HL1].e), "event");
    }
    check(gbh_event_record(c, 0, ev[EV_SETUP].e), "event");
    gbh_window_args w{b, E.d_win.as<uint32_t>(), nwin_, std::min(ncoop_, nwin_), E.d_i0.as<uint32_t>(),
                      nlarge_ ? E.d_hits.as<uint32_t>() : nullptr, E.d_hit_n.as<uint32_t>(), cap_, 0};
    check(gbh_launch_window(c, 0, &w), "window sieve");
    check(gbh_event_record(c, 0, ev[EV_WINDOW].e), "event");
    if (verify_every_ && nlarge_ && check_and_count)
        for (uint32_t s = 0; s < nseg; ++s)
            if (batch[s].index % verify_every_ == 0) enqueue_verify(s, b, out);
    if (check_and_count) {
        uint32_t kinds = 0;
        for (const SegJob& j : batch) kinds |= j.sampled ? 2u : 1u;
        gbh_check_args ca{b, E.ka, pcap_, E.d_coff.as<uint32_t>(), E.d_coffa.as<uint32_t>(),
                          E.d_pending.as<uint32_t>(), out, kinds, 0};
        check(gbh_launch_check_a(c, 0, &ca), "check phase A");
        gbh_checkb_args cb{b, E.ka, CHECK_TAIL, E.d_coff.as<uint32_t>(), E.d_pending.as<uint32_t>(), pcap_,
                           ridges_ ? ridges_->thresh() : 0xFFFFFFFFu, out};
        check(gbh_launch_check_b(c, 0, &cb), "check phase B");
    }
    check(gbh_event_record(c, 0, ev[EV_CHECK].e), "event");
    if (check_and_count && nwords) {
        gbh_count_args ct{b, E.d_cs.as<gbh_countseg>(), nwords, swords_, E.d_scratch.as<uint32_t>(), out};
        check(gbh_launch_count(c, 0, &ct), "count");
    }
    check(gbh_event_record(c, 0, ev[EV_COUNT].e), "event");
    check(gbh_copy_d2h(c, 0, E.h_out[slot].p, out, sizeof(gbh_segout) * nseg), "copy outputs");
    check(gbh_event_record(c, 0, ev[EV_END].e), "event");
}

// The GB_BUCKET_VERIFY rebuild of segment s (the reference's
// enqueue_bucket_verify, src/verifier.cpp:2059-2085): window primes as in
// production, every large prime by the wheel walk (no bitmap), then a word
// compare with the production map into out[s].vdiff.
void DenseRun::enqueue_verify(uint32_t s, const gbh_batch& b, gbh_segout* out)
{
    gbh_ctx* c = E.c;
    const gbh_batch vb{E.d_geoms.as<gbh_geom>() + s, 1, cw_, wbits_, b.nwin, E.d_vmap.as<uint64_t>()};
    if (nwin_) {
        const gbh_setup_args a{vb, E.d_win.as<uint32_t>(), nwin_, 0, E.d_i0.as<uint32_t>()};
        check(gbh_launch_setup(c, 0, &a), "verify setup");
    }
    check(gbh_zero(c, 0, E.d_vhit_n.p, size_t(8) * b.nwin * 4), "verify zero");
    const gbh_harvest_args h{vb, E.d_large.as<uint32_t>(), E.d_vchunks.as<uint32_t>(), nvchunks_, cap_v_, UINT32_MAX, 0,
                             nullptr, 0, 0, E.d_vhits.as<uint32_t>(), E.d_vhit_n.as<uint32_t>(), &out[s].err};
    check(gbh_launch_harvest(c, 0, &h), "verify harvest");
    const gbh_window_args w{vb, E.d_win.as<uint32_t>(), nwin_, std::min(ncoop_, nwin_), E.d_i0.as<uint32_t>(),
                            E.d_vhits.as<uint32_t>(), E.d_vhit_n.as<uint32_t>(), cap_v_, 0};
    check(gbh_launch_window(c, 0, &w), "verify window");
    const gbh_mapcmp_args m{E.d_map.as<uint64_t>() + size_t(s) * 8 * cw_, E.d_vmap.as<uint64_t>(),
                            E.d_geoms.as<gbh_geom>() + s, cw_, 0, &out[s].vdiff};
    check(gbh_launch_mapcmp(c, 0, &m), "verify compare");
    ++verified_;
}

void DenseRun::collect(std::vector<SegJob>& batch, int slot)
{
    gbh_ctx* c = E.c;
    auto& ev = E.ev[slot];
    check(gbh_event_sync(c, ev[EV_END].e), "batch sync");
    const double tl = elapsed_s(c, ev[EV_BEGIN], ev[EV_SETUP]);
    const double ts = elapsed_s(c, ev[EV_SETUP], ev[EV_WINDOW]);
    if (harvested_[slot]) {
        rs_.harvest_kernel_seconds += elapsed_s(c, ev[EV_HL0], ev[EV_HL1]);
        ++rs_.harvest_launches;
    }
    rs_.window_kernel_seconds += ts;
    ++rs_.window_launches;
    const double tc = elapsed_s(c, ev[EV_WINDOW], ev[EV_CHECK]);
    const double tn = elapsed_s(c, ev[EV_CHECK], ev[EV_COUNT]);
    t_sieve_ += ts; t_large_ += tl; t_check_ += tc; t_count_ += tn;
    const gbh_segout* outs = E.h_out[slot].as<gbh_segout>();
    if (outs[0].err & 1u) throw std::runtime_error("hit-list bucket overflow (raise the capacity)");
    if (outs[0].err & 4u) throw std::runtime_error("hit lists: a bitmap prime's cube is not above the batch's q range");
    if (outs[0].err & 8u) throw std::runtime_error("hit lists: a multiplier outside the multiplier bitmap");
    double share[4] = {ts / batch.size(), tl / batch.size(), tc / batch.size(), tn / batch.size()};
    for (size_t s = 0; s < batch.size(); ++s) collect_segment(batch[s], outs[s], share);
    if (phase_trace_)
        std::fprintf(stderr, "batch of %zu segments from %llu: harvest+setup %.3f ms window %.3f ms check %.3f ms "
                     "count %.3f ms\n", batch.size(), (unsigned long long)batch[0].sg.seg_start, tl * 1e3, ts * 1e3,
                     tc * 1e3, tn * 1e3);
}

void DenseRun::collect_segment(const SegJob& j, const gbh_segout& o, double share[4])
{
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
        if (p > 2 && gs_) gs_->add_cpu(n, p, j.sampled);
        if (p > 2 && ridges_) ridges_->consider(n, p, j.sg.seg_start);
    }
    rs_.leftovers += o.left_n;
    if (ridges_) ridges_->merge(o.ridge, o.ridge_n, GBH_RIDGE_CAP, j.sg.seg_start);
    if (gs_) {
        const size_t np = E.cand.primes.size();
        std::vector<uint64_t> h(15 * np, 0);
        for (uint32_t cr = 0; cr < 15; ++cr)
            for (size_t k = 0; k < E.cand.slot_pi[cr].size(); ++k)
                h[cr * np + E.cand.slot_pi[cr][k]] += o.hist[cr * GBH_NCAND + k];
        gs_->add_hist(j.sampled, h);
        gs_->add_first(j.sg.seg_start, o.first);
        if (gaps_on_) {
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
            gs_->add_gap_segment(g);
        }
        gs_->add_counts(o.cnt, o.twins);
        if (j.sampled) sampled_total_ += j.sg.seg_evens;
        evens_done_ += j.sg.seg_evens;
        gs_->write_live(start_, limit_, evens_done_, total_evens_, since(t_start_), false);
    }
    if (verbose_)
        std::printf("  seg %llu: [%llu, %llu] evens=%llu n_left=%u\n", (unsigned long long)j.index,
                    (unsigned long long)j.sg.seg_start, (unsigned long long)j.sg.seg_end,
                    (unsigned long long)j.sg.seg_evens, o.left_n);
    if (collect_) {
        SegmentStats st;
        st.seg_start = j.sg.seg_start;
        st.seg_end = j.sg.seg_end;
        st.evens = j.sg.seg_evens;
        st.tiles = uint64_t((j.cg.n[0] + wbits_ - 1) / wbits_) * 8;
        st.large_items = nlarge_;
        st.sieve_split = split_;
        st.mask_count = E.win_lo;
        st.leftovers = o.left_n;
        st.sieve_seconds = share[0];
        st.large_seconds = share[1];
        st.check_seconds = share[2] + share[3];
        rs_.per_segment.push_back(st);
    }
}

RunStats DenseRun::run()
{
    t_start_ = clock_::now();
    phase_trace_ = env::present("GB_PHASE");
    resolve_range();
    const double t_res = since(t_start_);
    configure_stats();
    configure_ridges();
    const double t_cfg = since(t_start_);
    plan_buffers();
    const auto t_gpu = clock_::now();
    rs_.host_plan_seconds = since(t_start_);
    if (phase_trace_)
        std::fprintf(stderr, "host: resolve %.3f ms, statistics set-up %.3f ms, plan %.3f ms\n", t_res * 1e3,
                     (t_cfg - t_res) * 1e3, (rs_.host_plan_seconds - t_cfg) * 1e3);
    std::vector<SegJob> batches[2];
    int cur = 0;
    bool pending_batch = false;
    uint64_t k = 0;
    bool done = false;
    uint64_t seg_start = start_;
    while (!done) {
        std::vector<SegJob>& bt = batches[cur];
        bt.clear();
        while (bt.size() < kb_ && !done) {
            SegJob j;
            j.sg = segment_geometry(seg_start, limit_, span_);
            j.index = k;
            j.sampled = stats_on_ && k % sample_ == 0;
            j.walked = gaps_on_ && (j.sampled || gaps_exact_);
            j.counted = stats_on_;
            bt.push_back(j);
            ++k;
            rs_.total_large_items += nlarge_;
            if (j.sg.seg_end == limit_) done = true;
            else seg_start += span_;
        }
        enqueue(bt, cur, true);
        for (const SegJob& j : bt) rs_.total_tiles += uint64_t((j.cg.n[0] + wbits_ - 1) / wbits_) * 8;
        if (pending_batch) collect(batches[cur ^ 1], cur ^ 1);
        pending_batch = true;
        cur ^= 1;
    }
    if (pending_batch) collect(batches[cur ^ 1], cur ^ 1);
    rs_.segments = k;
    rs_.gpu_seconds = since(t_gpu);
    rs_.init_seconds = since(t_start_) - rs_.gpu_seconds;
    rs_.total_seconds = rs_.init_seconds + rs_.gpu_seconds;
    rs_.evens_per_second = double(total_evens_) / rs_.gpu_seconds;
    rs_.sieve_seconds = t_sieve_;
    rs_.large_seconds = t_large_;
    rs_.verify_seconds = t_check_ + t_count_;
    rs_.count_seconds = t_count_;
    rs_.sync_seconds = std::max(0.0, rs_.gpu_seconds - t_sieve_ - t_large_ - t_check_ - t_count_);
    rs_.start = start_;
    if (verify_every_ && nlarge_)
        std::fprintf(stderr, "bucket verify: %llu segment maps rebuilt without the multiplier bitmap, identical\n",
                     (unsigned long long)verified_);
    if (ridges_ && ridges_->active()) {
        ridges_->save(ridge_path_);
        rs_.ridge_min = ridges_->floor();
        rs_.ridge_found = ridges_->appends();
        rs_.ridge_dropped = ridges_->dropped();
        rs_.ridge_top = ridges_->top();
        if (ridges_->overflowed())
            std::fprintf(stderr, "warning: a segment recorded more than %u ridges at or above the threshold; the "
                         "excess was not considered\n", GBH_RIDGE_CAP);
    }
    if (gs_) {
        gs_->finish_run(start_, limit_, total_evens_);
        rs_.stats_on = true;
        rs_.stats_pmax = gs_->run_pmax();
        rs_.stats_primes = gs_->run_primes();
        rs_.stats_twins = gs_->run_twins();
        rs_.stats_sampled = gs_->evens_sampled();
        rs_.stats_tail = gs_->run_tail();
        gs_->write_live(start_, limit_, total_evens_, total_evens_, since(t_start_), true);
        rs_.stats_line = gs_->log_line(VERSION);
        const auto tc = clock_::now();
        gs_->commit(VERSION, sampled_total_);
        if (phase_trace_) std::fprintf(stderr, "host: statistics commit %.3f ms\n", since(tc) * 1e3);
    }
    return rs_;
}

uint32_t DenseRun::map_of(uint64_t seg_start, std::vector<uint64_t>& words, uint64_t Q[8], uint32_t nc[8])
{
    resolve_range();
    plan_buffers();
    std::vector<SegJob> bt(1);
    bt[0].sg = segment_geometry(seg_start, limit_, span_);
    enqueue(bt, 0, false);
    check(gbh_event_sync(E.c, E.ev[0][EV_END].e), "sync");
    if (E.h_out[0].as<gbh_segout>()[0].err & 13u) throw std::runtime_error("hit-list device error");
    words.assign(size_t(8) * cw_, 0);
    check(gbh_copy_d2h(E.c, 0, words.data(), E.d_map.p, words.size() * 8), "copy map");
    check(gbh_stream_sync(E.c, 0), "sync");
    for (int c = 0; c < 8; ++c) { Q[c] = bt[0].cg.Q[c]; nc[c] = uint32_t(bt[0].cg.n[c]); }
    return cw_;
}

}  // namespace

Engine::Engine(uint64_t max_limit, uint64_t max_span, const char* prime_cache)
    : s_(std::make_unique<State>(max_limit, max_span, prime_cache))
{
}
Engine::~Engine() = default;
double Engine::init_seconds() const { return s_->init_s; }
double Engine::prepare(uint64_t limit, uint64_t start_even)
{
    DenseRun r(*s_, limit, s_->max_span, false, start_even, 0, false);
    return r.prepare();
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
    DenseRun r(*s_, limit, segment_even, verbose, start_even, sieve_split_arg, collect_segments);
    return r.run();
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
    DenseRun r(*e.state(), limit, span, false, seg_start, 0, false);
    return r.map_of(seg_start, words, Q, nc);
}

}  // namespace gb
