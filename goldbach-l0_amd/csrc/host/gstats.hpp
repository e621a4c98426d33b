#pragma once
// Exact Goldbach statistics of a run and of a campaign.
//
// What is measured (every field exact):
//   * p(n) of every even with p(n) >= tail_t (3000), by prime, and the first
//     even of the run with each such p;
//   * p(n) of every even of the sampled segments (segment index % sample
//     == 0), by even class n mod 30 / 2 and prime;
//   * primes and twin pairs (q, q + 2) with q in [start + 1, limit - 1]
//     (twins: q + 2 <= limit - 1), by residue mod 30;
//   * the gap q - prev(q) of every prime q in [start + 1, limit - 1]: count,
//     sum, every gap >= 700 with its first lower prime, the largest gap;
//     histogram and moments over the walked segments;
//   * evens whose p(n) the CPU fallback found beyond the candidate list.
//
// Files, in the reference's formats so an existing campaign and its scripts
// keep working (include/gstats.hpp of the reference: log line 913-969,
// state 1267-1315, gap state 1014-1048, records 824-868): the GBSTATS v2
// log line (GBPSAMP / GBGAPS rows for a range already covered), the state
// file v2 and its ".gaps" companion v1, the records file and the live JSON
// snapshot.
//
// Design: a run and a campaign are the same additive Tally (the run's is
// absorbed into the campaign's on commit); the state file is the Tally's
// text form, read back by a table of line rules; gaps live in a GapBook
// with the same absorb step; records (RECORD / BEYOND / SP events) go
// through one sink that re-proves them by an independent CPU scan.
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace gb {

using u128 = unsigned __int128;
std::string u128_str(u128 v);

inline constexpr uint32_t GAP_TAIL = 700;              // gaps listed one by one from here (GAPS30_TAIL)
inline constexpr uint32_t GAP_BINS = 1024;             // histogram bins g / 2 (GAPS30_HBINS)
inline constexpr uint64_t STATS_MIN_START = 1000000;   // statistics only from this start on
inline constexpr uint64_t RECORD_P = 9781;             // the largest p(n) known below 4e18

/// Prime gaps of a run or a campaign.
struct GapBook {
    struct Tail {
        uint64_t count = 0, first_lower = 0;
    };
    uint64_t count = 0;            // gaps (one per prime of the range)
    u128 span = 0;                 // their sum: last prime - prime before the first
    std::map<uint32_t, Tail> tail; // g >= GAP_TAIL
    uint64_t top_g = 0, top_lower = 0;   // the largest gap seen and its first lower prime
    bool all_walked = true;        // every segment walked: top_g is the largest gap
    uint64_t walk_count = 0;       // gaps of the walked segments
    u128 moment[4] = {};           // their sums of g, g^2, g^3, g^4
    std::vector<uint64_t> walk_hist = std::vector<uint64_t>(GAP_BINS, 0);   // by g / 2

    /// One gap of a walked segment added (+1) or taken out (-1).
    void walk(uint64_t g, int sign);
    /// One exactly known gap: a tail entry when g >= GAP_TAIL, and the top.
    void exact(uint64_t g, uint64_t lower);
    void offer_top(uint64_t g, uint64_t lower);
    void add_tail(uint32_t g, uint64_t n, uint64_t lower);
    bool top_exact() const { return top_g >= GAP_TAIL || all_walked; }
    void absorb(const GapBook& o);
};

/// One segment's counter outputs as the statistics consume them.
struct SegmentGaps {
    uint64_t q_first = 0;          // position origin
    uint64_t evens = 0;
    uint64_t primes = 0;           // counted primes of the segment
    uint32_t pfirst = ~0u, plast = 0;   // positions of the first / last counted prime
    const uint32_t* tail = nullptr;     // (upper position, g) pairs, inner gaps >= 700
    uint32_t ntail = 0;
    bool walked = false;
    uint64_t gmax = 0;             // largest inner gap: g << 32 | ~position (0: none)
    const uint32_t* hist = nullptr;     // inner gaps by g / 2 (GAP_BINS + 1 bins)
};

/// The additive statistics of a run or a campaign (per candidate prime
/// index i and even class cr).
struct Tally {
    uint64_t evens = 0, sampled = 0;
    uint64_t primes[8] = {}, twins[3] = {};
    std::vector<uint64_t> hist;        // [cr][i]: p(n) of the sampled evens
    std::vector<uint64_t> tail_cls;    // [cr][i]: p(n) >= tail_t, sampled evens
    std::vector<uint64_t> tail_flat;   // [i]:     p(n) >= tail_t, the other evens
    std::vector<uint64_t> first;       // [i]: first even with p(n) = p_i (0: none)
    std::vector<std::pair<uint64_t, uint64_t>> beyond;   // (n, p): p(n) not among the candidates
    std::vector<std::pair<uint64_t, uint64_t>> ranges;   // covered [start, limit], merged
    bool zero = false;                 // the four count vectors are known to be all zero (clear() skips them)

    /// Zeroed vectors for np candidates; by_class false leaves hist and
    /// tail_cls empty (size_by_class or absorb sizes them when needed).
    void size_for(size_t np, bool by_class);
    void size_by_class(size_t np);
    void clear();
    /// Sums, first-occurrence minima, the union of ranges.
    void absorb(const Tally& run);
};

class GoldbachStats {
public:
    /// @param cands  odd candidate primes 3 .. 9781 ascending.
    /// @param tail_t every even with p(n) >= tail_t is recorded.
    /// @param campaign true: the campaign files (GB_STATS_STATE, default
    ///        goldbach_stats.state, its ".gaps" file, GB_STATS_LOG, ...) are
    ///        read here (a damaged file fails before the run) and written by
    ///        commit(); false: a run of its own (batched and served ranges),
    ///        no file is touched and memory_commit() returns the state text a
    ///        fresh campaign would hold.
    GoldbachStats(const std::vector<uint32_t>& cands, uint32_t tail_t, bool campaign = true);
    GoldbachStats(const GoldbachStats&) = delete;
    GoldbachStats& operator=(const GoldbachStats&) = delete;
    ~GoldbachStats();

    /// The run starts at @p start: first occurrences the campaign holds at
    /// or below it are not reported again.
    void set_run_start(uint64_t start);
    bool wanted(uint32_t pi) const { return want_[pi] != 0; }

    /// A segment's p(n) counts from the device's per-slot table:
    /// counts[cr * stride + k] evens of class cr resolved by candidate slot
    /// k, slot_pi[cr][k] its prime index. Sampled: all count, else only
    /// p >= tail_t. @return evens added to the sampled histogram.
    uint64_t add_hist_slots(bool sampled, const uint32_t* counts, size_t stride,
                            const std::vector<uint16_t>* slot_pi);
    /// First even index (from seg_start) by prime index (~0u: none).
    void add_first(uint64_t seg_start, const uint32_t* first_by_prime);
    void add_counts(const uint32_t cnt[8], const uint32_t twins[3]);
    void begin_gaps(uint64_t start);
    /// @throws std::runtime_error on a failed cross-check of the walk.
    void add_gap_segment(const SegmentGaps& s);
    /// A CPU-resolved even (a leftover); p may exceed 9781.
    void add_cpu(uint64_t n, uint64_t p, bool sampled);
    /// Edge corrections: the device counts from start - 5 on; the run
    /// reports q in [start + 1, limit - 1].
    void finish_run(uint64_t start, uint64_t limit, uint64_t evens);
    /// Merges the run into the campaign state and appends its log line
    /// (for a range that overlaps a covered one: only the sampled-histogram
    /// and gap rows the campaign lacks). @p run_line: log_line(version),
    /// reused when the campaign's first occurrences did not change meanwhile.
    void commit(const char* version, uint64_t sampled_total, const std::string* run_line = nullptr);
    std::string memory_commit(uint64_t sampled_total);
    std::string log_line(const char* version) const;
    void write_live(uint64_t start, uint64_t limit, uint64_t done, uint64_t total, double elapsed, bool force);

    uint64_t run_primes() const;
    uint64_t run_twins() const { return run_.twins[0] + run_.twins[1] + run_.twins[2]; }
    uint64_t run_tail() const;
    uint64_t run_pmax() const;
    uint64_t evens_sampled() const { return run_.sampled; }
    const GapBook& run_gaps() const { return gaps_; }
    bool gaps_started() const { return gaps_from_ != ~0ull; }

private:
    struct GapFile {               // the ".gaps" companion of the state file
        GapBook book;
        uint64_t evens = 0, walked = 0;
        std::vector<std::pair<uint64_t, uint64_t>> ranges;
    };
    void count_hit(bool sampled, uint32_t cr, uint32_t pi, uint64_t c);
    void note_first(uint32_t pi, uint64_t n);
    void announce(uint64_t n, uint64_t p, const char* kind, const char* what);
    std::string gap_fields() const;
    std::vector<uint64_t> per_prime(const std::vector<uint64_t>& by_class, const std::vector<uint64_t>* flat) const;
    void read_state(Tally& into) const;
    std::string state_text(const Tally& t) const;
    GapFile read_gap_file() const;
    bool commit_gap_file();
    bool log_has_gap_row(uint64_t a, uint64_t b) const;
    void log_append(const std::string& line) const;

    const std::vector<uint32_t>& cands_;
    const uint32_t tail_t_;
    size_t np_ = 0, tail0_ = 0;        // candidates; the first with p >= tail_t
    std::vector<int32_t> index_of_;    // p / 2 -> prime index (-1: not a candidate)
    std::string state_path_, log_path_, rec_path_, live_path_;
    bool campaign_ = true;
    Tally camp_, run_;
    GapBook gaps_;
    uint64_t gaps_from_ = ~0ull;       // the run start the gaps were begun for
    uint64_t last_prime_ = 0;          // the last prime before the next segment's first
    uint64_t before_start_ = 0;        // prev_prime(start - 3) of the gap run (set with last_prime_'s first value)
    bool last_walked_ = false;
    uint64_t walked_evens_ = 0;
    uint64_t start_ = 0, limit_ = 0;
    std::vector<uint8_t> want_;
    std::vector<std::pair<uint64_t, uint64_t>> events_;   // (n, p) announced this run
    long long live_stamp_ = 0;
};

}  // namespace gb
