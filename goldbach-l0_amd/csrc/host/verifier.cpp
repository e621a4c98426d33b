// GPU verification driver (contracts in verifier.hpp). The run is the
// reference's DenseRun (src/verifier.cpp:2615-2642) restated for the MI355X
// pipeline: range resolution and segmentation exactly as the reference
// (statistics depend on them), base primes, then a loop that enqueues
// segment k on one in-order HIP stream while the host collects segment k-1
// (two output slots), fail-closed coverage accounting, the leftover CPU
// fallback and the exact statistics.
#include "verifier.hpp"

#include <algorithm>
#include <chrono>
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
#include "primes.hpp"

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
    Buf(gbh_ctx* ctx, size_t n, int k) : c(ctx), kind(k), bytes(n)
    {
        check(gbh_alloc(c, std::max<size_t>(n, 256), k, &p), "gbh_alloc");
    }
    ~Buf() { if (p) (void)gbh_free(c, p, kind); }
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf& operator=(Buf&& o) noexcept
    {
        std::swap(c, o.c); std::swap(p, o.p); std::swap(kind, o.kind); std::swap(bytes, o.bytes);
        return *this;
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

enum { EV_BEGIN, EV_WINDOW, EV_LARGE, EV_CHECK, EV_COUNT, EV_END, EV_N };

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

}  // namespace

// ---------------------------------------------------------------- Engine state

struct Engine::State {
    clock_::time_point t0;
    uint64_t max_limit = 0, max_span = 0;
    std::vector<uint32_t> base;          // odd primes <= sqrt(max_limit) + 16
    uint32_t win_lo = 0, win_hi = 0;     // window primes base[win_lo, win_hi), large base[win_hi, ...)
    uint32_t wbits = 0, ka = 0, ncoop = 0;
    double init_s = 0, cache_s = 0, module_s = 0, upload_s = 0;
    std::optional<Ctx> ctx;
    gbh_ctx* c = nullptr;
    Buf d_win, d_large, d_i0, d_map, d_coff, d_pending, d_scratch;
    Buf d_out[2], h_out[2];
    uint64_t cw_cap = 0;
    uint32_t pending_cap = 0, scratch_words = 0;
    Event ev[2][EV_N];
    Candidates cand;

    State(uint64_t max_limit_, uint64_t max_span_)
    {
        t0 = clock_::now();
        require_supported_limit(max_limit_);
        max_limit = max_limit_;
        max_span = std::max<uint64_t>(max_span_ + (max_span_ & 1), 4);
        if (max_span > (uint64_t(1) << 35)) throw std::runtime_error("segment span above 2^35 integers");
        const uint32_t kb = env::u32("GB_WIN_KB", WIN_KB_DEFAULT, 4, 128);
        if (kb & (kb - 1)) throw std::runtime_error("GB_WIN_KB must be a power of two");
        wbits = kb * 8192;
        ka = env::u32("GB_KA", KA_DEFAULT, 1, 128);
        // base primes (the reference loads them from its prime cache;
        // here a segmented sieve on the host threads)
        const auto tc = clock_::now();
        base = base_primes(std::max<uint32_t>(base_prime_bound(max_limit), 64), host_threads());
        cache_s = since(tc);
        win_lo = uint32_t(std::lower_bound(base.begin(), base.end(), 67u) - base.begin());
        const uint64_t win_p = env::u64("GB_WIN_P", WIN_P_DEFAULT, 67, UINT32_MAX);
        win_hi = std::max(win_lo, uint32_t(std::lower_bound(base.begin(), base.end(), win_p) - base.begin()));
        // window primes with at least two marks per lane per window are
        // marked by the whole work-group
        while (win_lo + ncoop < win_hi && base[win_lo + ncoop] < wbits / 128) ++ncoop;
        const auto tm = clock_::now();
        ctx.emplace();
        c = ctx->c;
        module_s = since(tm);
        const auto tu = clock_::now();
        const uint32_t nwin = win_hi - win_lo, nlarge = uint32_t(base.size()) - win_hi;
        d_win = Buf(c, size_t(nwin) * 4, 0);
        d_large = Buf(c, size_t(nlarge) * 4, 0);
        if (nwin) check(gbh_copy_h2d(c, 0, d_win.p, base.data() + win_lo, size_t(nwin) * 4), "upload window primes");
        if (nlarge) check(gbh_copy_h2d(c, 0, d_large.p, base.data() + win_hi, size_t(nlarge) * 4), "upload large primes");
        d_i0 = Buf(c, size_t(nwin) * 8 * 4, 0);
        // the largest segment: the merged last one spans below 2 x span
        const uint64_t qodds_max = (2 * max_span + MAX_GB_PRIME) / 2 + 2;
        const uint64_t nc_max = (qodds_max + 14) / 15 + 1;
        const uint64_t nwin_max = (nc_max + wbits - 1) / wbits;
        cw_cap = nwin_max * (wbits / 64) + 16;
        d_map = Buf(c, size_t(8) * cw_cap * 8, 0);
        check(gbh_zero(c, 0, d_map.p, size_t(8) * cw_cap * 8), "zero map");
        d_coff = Buf(c, size_t(15) * GBH_NCAND * 4, 0);
        pending_cap = uint32_t(std::min<uint64_t>((2 * max_span + 4) / 4 + 4096, UINT32_MAX / 2));
        d_pending = Buf(c, size_t(pending_cap) * 4, 0);
        scratch_words = uint32_t(2 * (nc_max / 64 + 4));
        d_scratch = Buf(c, size_t(scratch_words) * 4, 0);
        for (int s = 0; s < 2; ++s) {
            d_out[s] = Buf(c, sizeof(gbh_segout), 0);
            h_out[s] = Buf(c, sizeof(gbh_segout), 1);
            for (auto& e : ev[s]) e.create(c);
        }
        check(gbh_stream_sync(c, 0), "upload sync");
        upload_s = since(tu);
        init_s = since(t0);
    }
};

namespace {

struct SegJob {
    SegmentGeometry sg;
    ClassGeo cg;
    gbh_geom geom{};
    uint64_t index = 0;
    uint64_t n_gpu = 0;
    uint64_t gpu_evens = 0;
    int slot = 0;
    bool sampled = false, walked = false, counted = false;
    clock_::time_point t_wall;
};

// One run over [start, limit] on an engine's resident state (the
// reference's DenseRun phases: resolve_range, statistics set-up, the
// two-slot segment loop, finish and commit).
class DenseRun {
public:
    DenseRun(Engine::State& e, uint64_t limit, uint64_t segment_even, bool verbose, uint64_t start_even,
             uint32_t split_arg, bool collect)
        : E(e), limit_(limit), span_(segment_even), verbose_(verbose), start_(start_even), split_arg_(split_arg),
          collect_(collect)
    {
    }
    RunStats run();

private:
    void resolve_range();
    void configure_stats();
    void enqueue(SegJob& j);
    void collect(SegJob& j);

    Engine::State& E;
    uint64_t limit_, span_;
    bool verbose_;
    uint64_t start_;
    uint32_t split_arg_;
    bool collect_;
    RunStats rs_;
    uint64_t total_evens_ = 0;
    clock_::time_point t_start_;
    uint32_t nwin_ = 0, nlarge_ = 0, split_ = 0;   // this run's window / large prime counts
    bool stats_on_ = false, gaps_on_ = false, gaps_exact_ = false;
    uint32_t sample_ = 256;
    std::optional<GoldbachStats> gs_;
    uint64_t sampled_total_ = 0, evens_done_ = 0;
    bool phase_trace_ = false;
    double t_sieve_ = 0, t_large_ = 0, t_check_ = 0, t_count_ = 0;
};

void DenseRun::resolve_range()
{
    if (limit_ < 4) throw std::runtime_error("limit < 4: no even n >= 4 to verify");
    refuse_removed_knobs();
    require_supported_limit(limit_);
    limit_ &= ~uint64_t(1);
    rs_.limit = limit_;
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
    // this run's primes: those <= sqrt(limit) + 16
    const uint32_t bound = std::max<uint32_t>(base_prime_bound(limit_), 64);
    const uint32_t nb = uint32_t(std::upper_bound(E.base.begin(), E.base.end(), bound) - E.base.begin());
    uint32_t hi = std::min(E.win_hi, nb);
    if (split_arg_ > 0) {          // --split: the first large prime's index among the base primes
        if (split_arg_ < E.win_lo || split_arg_ > E.win_hi)
            throw std::runtime_error("--split must lie in [" + std::to_string(E.win_lo) + ", " +
                                     std::to_string(E.win_hi) + "] on this engine");
        hi = std::min(split_arg_, nb);
    }
    nwin_ = hi > E.win_lo ? hi - E.win_lo : 0;
    split_ = E.win_lo + nwin_;
    nlarge_ = nb > E.win_hi ? nb - E.win_hi : 0;
    if (split_ < E.win_hi && nb > split_) {
        // a lowered split: primes [split, win_hi) would go to the large stage,
        // which takes the device list from win_hi on; keep them in the windows
        nwin_ = std::min(E.win_hi, nb) - E.win_lo;
        split_ = E.win_lo + nwin_;
    }
    rs_.split_used = split_;
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

void DenseRun::enqueue(SegJob& j)
{
    gbh_ctx* c = E.c;
    j.cg = class_geometry(j.sg.q_first, j.sg.q_odds);
    j.geom = make_geom(j.sg, j.cg, E.wbits);
    if (j.geom.cw > E.cw_cap) throw std::runtime_error("segment class stride above the allocation");
    // evens below q_first + 9780 (only at the very bottom of the range) have
    // candidate q outside the map: the CPU resolves them
    uint64_t n_gpu = std::max<uint64_t>(j.sg.seg_start, j.sg.q_first + 9780);
    n_gpu += n_gpu & 1;
    j.n_gpu = n_gpu;
    j.gpu_evens = n_gpu <= j.sg.seg_end ? (j.sg.seg_end - n_gpu) / 2 + 1 : 0;
    const int s = j.slot;
    gbh_segout* out = E.d_out[s].as<gbh_segout>();
    check(gbh_event_record(c, 0, E.ev[s][EV_BEGIN].e), "event");
    check(gbh_zero(c, 0, out, sizeof(gbh_segout)), "zero outputs");
    check(gbh_fill32(c, 0, out->first, GBH_NCAND, ~0u), "fill first");
    check(gbh_fill32(c, 0, &out->pfirst, 1, ~0u), "fill pfirst");
    if (nwin_) {
        gbh_setup_args a{&j.geom, E.d_win.as<uint32_t>(), nwin_, E.d_i0.as<uint32_t>()};
        check(gbh_launch_setup(c, 0, &a), "setup");
    }
    gbh_window_args w{&j.geom, E.d_win.as<uint32_t>(), nwin_, std::min(E.ncoop, nwin_), E.d_i0.as<uint32_t>(),
                      E.d_map.as<uint64_t>()};
    check(gbh_launch_window(c, 0, &w), "window sieve");
    check(gbh_event_record(c, 0, E.ev[s][EV_WINDOW].e), "event");
    if (nlarge_) {
        gbh_large_args l{&j.geom, E.d_large.as<uint32_t>(), nlarge_, 0, E.d_map.as<uint64_t>()};
        check(gbh_launch_large(c, 0, &l), "large primes");
    }
    check(gbh_event_record(c, 0, E.ev[s][EV_LARGE].e), "event");
    if (j.gpu_evens) {
        gbh_check_args ca{&j.geom, E.d_map.as<uint64_t>(), n_gpu, E.ka, j.sampled ? 1u : 0u, E.d_coff.as<uint32_t>(),
                          E.d_pending.as<uint32_t>(), E.pending_cap, 0, out};
        check(gbh_launch_check_a(c, 0, &ca), "check phase A");
        gbh_checkb_args cb{&j.geom, E.d_map.as<uint64_t>(), n_gpu, E.ka, CHECK_TAIL, E.d_coff.as<uint32_t>(),
                           E.d_pending.as<uint32_t>(), E.pending_cap, 0, out};
        check(gbh_launch_check_b(c, 0, &cb), "check phase B");
    }
    check(gbh_event_record(c, 0, E.ev[s][EV_CHECK].e), "event");
    if (j.counted) {
        gbh_count_args ct = make_count_args(j.sg, j.cg);
        ct.geom = &j.geom;
        ct.map = E.d_map.as<uint64_t>();
        ct.scratch = E.d_scratch.as<uint32_t>();
        ct.scratch_words = E.scratch_words;
        ct.out = out;
        check(gbh_launch_count(c, 0, &ct), "count");
    }
    check(gbh_event_record(c, 0, E.ev[s][EV_COUNT].e), "event");
    check(gbh_copy_d2h(c, 0, E.h_out[s].p, out, sizeof(gbh_segout)), "copy outputs");
    check(gbh_event_record(c, 0, E.ev[s][EV_END].e), "event");
}

void DenseRun::collect(SegJob& j)
{
    gbh_ctx* c = E.c;
    const int s = j.slot;
    check(gbh_event_sync(c, E.ev[s][EV_END].e), "segment sync");
    const gbh_segout& o = *E.h_out[s].as<gbh_segout>();
    const double ts = elapsed_s(c, E.ev[s][EV_BEGIN], E.ev[s][EV_WINDOW]);
    const double tl = elapsed_s(c, E.ev[s][EV_WINDOW], E.ev[s][EV_LARGE]);
    const double tc = elapsed_s(c, E.ev[s][EV_LARGE], E.ev[s][EV_CHECK]);
    const double tn = elapsed_s(c, E.ev[s][EV_CHECK], E.ev[s][EV_COUNT]);
    t_sieve_ += ts; t_large_ += tl; t_check_ += tc; t_count_ += tn;
    const std::string seg = " in segment [" + std::to_string(j.sg.seg_start) + ", " + std::to_string(j.sg.seg_end) + "]";
    if (o.err) throw std::runtime_error("device error flags " + std::to_string(o.err) + seg);
    if (o.pending_n > E.pending_cap)
        throw std::runtime_error("pending overflow (" + std::to_string(o.pending_n) + ")" + seg);
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
    for (uint64_t n = j.sg.seg_start; n < j.n_gpu && n <= j.sg.seg_end; n += 2) cpu_n.push_back(n);
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
    }
    rs_.leftovers += o.left_n;
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
    if (phase_trace_)
        std::fprintf(stderr, "seg %llu [%llu, %llu]: window %.3f ms large %.3f ms check %.3f ms count %.3f ms pending %u\n",
                     (unsigned long long)j.index, (unsigned long long)j.sg.seg_start, (unsigned long long)j.sg.seg_end,
                     ts * 1e3, tl * 1e3, tc * 1e3, tn * 1e3, o.pending_n);
    if (verbose_)
        std::printf("  seg %llu: [%llu, %llu] evens=%llu n_left=%u\n", (unsigned long long)j.index,
                    (unsigned long long)j.sg.seg_start, (unsigned long long)j.sg.seg_end,
                    (unsigned long long)j.sg.seg_evens, o.left_n);
    if (collect_) {
        SegmentStats st;
        st.seg_start = j.sg.seg_start;
        st.seg_end = j.sg.seg_end;
        st.evens = j.sg.seg_evens;
        st.tiles = uint64_t(j.geom.nwin) * 8;
        st.large_items = nlarge_;
        st.sieve_split = split_;
        st.mask_count = E.win_lo;
        st.leftovers = o.left_n;
        st.sieve_seconds = ts;
        st.large_seconds = tl;
        st.check_seconds = tc + tn;
        st.wall_seconds = since(j.t_wall);
        rs_.per_segment.push_back(st);
    }
}

RunStats DenseRun::run()
{
    t_start_ = clock_::now();
    resolve_range();
    configure_stats();
    phase_trace_ = env::present("GB_PHASE");
    const auto t_gpu = clock_::now();
    SegJob jobs[2];
    SegJob* prev = nullptr;
    uint64_t k = 0;
    for (uint64_t seg_start = start_; seg_start <= limit_; seg_start += span_, ++k) {
        SegJob& j = jobs[k & 1];
        j = SegJob{};
        j.sg = segment_geometry(seg_start, limit_, span_);
        j.index = k;
        j.slot = int(k & 1);
        j.sampled = stats_on_ && k % sample_ == 0;
        j.walked = gaps_on_ && (j.sampled || gaps_exact_);
        j.counted = stats_on_;
        j.t_wall = clock_::now();
        enqueue(j);
        rs_.total_tiles += uint64_t(j.geom.nwin) * 8;
        rs_.total_large_items += nlarge_;
        if (prev) collect(*prev);
        prev = &j;
        if (j.sg.seg_end == limit_) { ++k; break; }
    }
    if (prev) collect(*prev);
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
        gs_->commit(VERSION, sampled_total_);
    }
    return rs_;
}

}  // namespace

Engine::Engine(uint64_t max_limit, uint64_t max_span) : s_(std::make_unique<State>(max_limit, max_span)) {}
Engine::~Engine() = default;
double Engine::init_seconds() const { return s_->init_s; }
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

RunStats run_gpu(uint64_t limit, uint64_t segment_even, const std::string& spv_path, bool verbose, uint64_t start_even,
                 uint32_t sieve_split_arg, const char* prime_cache_path, bool collect_segments)
{
    (void)spv_path;
    (void)prime_cache_path;
    refuse_removed_knobs();
    if (limit < 4) throw std::runtime_error("limit < 4: no even n >= 4 to verify");
    require_supported_limit(limit);
    // the engine is sized for this run: the span the range resolution keeps
    const uint64_t lim = limit & ~uint64_t(1);
    const uint64_t st = std::max<uint64_t>(start_even + (start_even & 1), 4);
    uint64_t span = std::max<uint64_t>(segment_even + (segment_even & 1), 2);
    if (st <= lim) span = std::min<uint64_t>(span, std::max<uint64_t>((((lim - st) / 2 + 1) & ~uint64_t(1)), 4));
    Engine e(lim, std::max<uint64_t>(span, 4));
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
                "  clock         : %d MHz\n  window default: %u KB LDS, window primes < %llu, phase A %u candidates\n",
                i.name, i.arch, i.cus, i.lds_per_block, double(i.hbm_bytes) / 1e9, i.clock_khz / 1000, WIN_KB_DEFAULT,
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
    const SegmentGeometry sg = segment_geometry(seg_start, limit, span);
    const ClassGeo cg = class_geometry(sg.q_first, sg.q_odds);
    const uint32_t kb = env::u32("GB_WIN_KB", WIN_KB_DEFAULT, 4, 128);
    const gbh_geom g = make_geom(sg, cg, kb * 8192);
    const uint32_t bound = std::max<uint32_t>(base_prime_bound(limit), 64);
    const std::vector<uint32_t> base = base_primes(bound, host_threads());
    const uint32_t lo = uint32_t(std::lower_bound(base.begin(), base.end(), 67u) - base.begin());
    const uint64_t win_p = env::u64("GB_WIN_P", WIN_P_DEFAULT, 67, UINT32_MAX);
    const uint32_t hi = std::max(lo, uint32_t(std::lower_bound(base.begin(), base.end(), win_p) - base.begin()));
    Ctx c;
    Buf dw(c.c, size_t(hi - lo) * 4, 0), dl(c.c, (base.size() - hi) * 4, 0), di0(c.c, size_t(hi - lo) * 32, 0);
    Buf dm(c.c, size_t(8) * g.cw * 8, 0);
    check(gbh_zero(c.c, 0, dm.p, size_t(8) * g.cw * 8), "zero");
    if (hi > lo) check(gbh_copy_h2d(c.c, 0, dw.p, base.data() + lo, size_t(hi - lo) * 4), "h2d");
    if (base.size() > hi) check(gbh_copy_h2d(c.c, 0, dl.p, base.data() + hi, (base.size() - hi) * 4), "h2d");
    uint32_t ncoop = 0;
    while (lo + ncoop < hi && base[lo + ncoop] < g.wbits / 128) ++ncoop;
    if (hi > lo) {
        gbh_setup_args a{&g, dw.as<uint32_t>(), hi - lo, di0.as<uint32_t>()};
        check(gbh_launch_setup(c.c, 0, &a), "setup");
    }
    gbh_window_args w{&g, dw.as<uint32_t>(), hi - lo, ncoop, di0.as<uint32_t>(), dm.as<uint64_t>()};
    check(gbh_launch_window(c.c, 0, &w), "window");
    if (base.size() > hi) {
        gbh_large_args l{&g, dl.as<uint32_t>(), uint32_t(base.size() - hi), 0, dm.as<uint64_t>()};
        check(gbh_launch_large(c.c, 0, &l), "large");
    }
    words.assign(size_t(8) * g.cw, 0);
    check(gbh_copy_d2h(c.c, 0, words.data(), dm.p, words.size() * 8), "d2h");
    check(gbh_stream_sync(c.c, 0), "sync");
    for (int k = 0; k < 8; ++k) { Q[k] = g.Q[k]; nc[k] = g.nc[k]; }
    return g.cw;
}

}  // namespace gb
