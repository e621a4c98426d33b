#pragma once
// Exact Goldbach statistics of a run and the campaign state: the GBSTATS v2
// log line, the state file v2 and its ".gaps" companion v1, the records
// file and the live snapshot. The formats, and the meaning of every field,
// are the reference's (include/gstats.hpp: GapStats 397-454, GoldbachStats
// 457-1367, log line 913-969, state 1267-1315, gap state 1014-1048), so a
// campaign's existing files and scripts keep working; the accumulation code
// is this build's own and is fed from this build's per-segment device
// outputs (gbh_segout).
//
// Field semantics (all exact):
//   * p(n) of every even with p(n) >= tail_t (3000), by prime, and the
//     first even of the run with each such p;
//   * p(n) of every even of the sampled segments (segment index % sample
//     == 0), by even class n mod 30 / 2 and prime;
//   * primes and twin pairs (q, q + 2) with q in [start + 1, limit - 1]
//     (twins: q + 2 <= limit - 1), by residue;
//   * gaps q - prev(q) for every prime q in [start + 1, limit - 1]: count,
//     sum, every gap >= 700 with its first lower prime, the largest gap;
//     histogram and moments on the walked segments;
//   * leftovers (p(n) not found among the candidates) by the CPU.
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace gb {

using u128 = unsigned __int128;
std::string u128_str(u128 v);

inline constexpr uint32_t GAP_TAIL = 700;      // gstats.hpp:74
inline constexpr uint32_t GAP_BINS = 1024;     // gstats.hpp:75
inline constexpr uint64_t STATS_MIN_START = 1000000;   // gstats.hpp:125
inline constexpr uint64_t RECORD_P = 9781;     // gstats.hpp:137

/// Gap statistics of a run or a campaign (the reference's GapStats).
struct GapTotals {
    uint64_t n = 0;
    u128 d1 = 0;
    std::map<uint32_t, std::pair<uint64_t, uint64_t>> tail;   // g -> (count, first lower prime)
    uint64_t gmax = 0, gmax_lo = 0;
    bool all_walked = true;
    uint64_t wn = 0;
    u128 wd[4] = {};
    std::vector<uint64_t> samp = std::vector<uint64_t>(GAP_BINS, 0);

    void walk_add(uint64_t g);
    void walk_remove(uint64_t g);
    void exact_gap(uint64_t g, uint64_t lower);
    void note_max(uint64_t g, uint64_t lower);
    void note_tail(uint32_t g, uint64_t count, uint64_t lower);
    bool gmax_exact() const { return gmax >= GAP_TAIL || all_walked; }
    void merge(const GapTotals& o);
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

class GoldbachStats {
public:
    /// @param cands  odd candidate primes 3 .. 9781 ascending.
    /// @param tail_t every even with p(n) >= tail_t is recorded.
    /// Loads the campaign state (GB_STATS_STATE, default goldbach_stats.state)
    /// and its gap state; a damaged file fails here, before the run.
    /// @param campaign false: a run of its own (batched and served ranges):
    ///        no campaign files are read or written, the run starts from a
    ///        fresh state and memory_commit() returns the state text.
    GoldbachStats(const std::vector<uint32_t>& cands, uint32_t tail_t, bool campaign = true);
    GoldbachStats(const GoldbachStats&) = default;
    ~GoldbachStats();

    void set_run_start(uint64_t start);
    /// Tail prime indices without a campaign first occurrence at or below the
    /// run start (the first occurrences the run reports).
    bool wanted(uint32_t pi) const { return want_[pi] != 0; }

    /// One segment's p(n) counts by even class cr and prime index
    /// ([15][ncand]); sampled: all of them count, else only p >= tail_t.
    /// @return evens counted into the sampled histogram.
    uint64_t add_hist(bool sampled, const std::vector<uint64_t>& by_class_prime);
    /// The same from the device's per-slot counts: counts[cr * stride + k]
    /// evens of class cr resolved by slot k, slot_pi[cr][k] its prime index.
    uint64_t add_hist_slots(bool sampled, const uint32_t* counts, size_t stride,
                            const std::vector<uint16_t>* slot_pi);
    /// First even index (from seg_start) by prime index (~0u: none).
    void add_first(uint64_t seg_start, const uint32_t* first_by_prime);
    void add_counts(const uint32_t cnt[8], const uint32_t twins[3]);
    void begin_gaps(uint64_t start);
    /// @throws std::runtime_error on a failed cross-check of the walk.
    void add_gap_segment(const SegmentGaps& s);
    /// CPU-resolved even (leftover); p may exceed 9781.
    void add_cpu(uint64_t n, uint64_t p, bool sampled);
    void finish_run(uint64_t start, uint64_t limit, uint64_t evens);
    /// Merges the run into the campaign state and appends its log line
    /// (nothing when the range overlaps a covered one).
    /// run_line: log_line(version) if the caller already has it (the run's
    /// data is unchanged since).
    void commit(const char* version, uint64_t sampled_total, const std::string* run_line = nullptr);
    /// A run without campaign files: the state file text a fresh campaign
    /// holds after committing this run alone (same checks as commit()).
    std::string memory_commit(uint64_t sampled_total);
    std::string log_line(const char* version) const;
    void write_live(uint64_t start, uint64_t limit, uint64_t done, uint64_t total, double elapsed, bool force);

    uint64_t run_primes() const;
    uint64_t run_twins() const { return twins_[0] + twins_[1] + twins_[2]; }
    uint64_t run_tail() const;
    uint64_t run_pmax() const;
    uint64_t evens_sampled() const { return evens_sampled_; }
    const GapTotals& run_gaps() const { return gaps_; }
    bool gaps_started() const { return gap_start_ != ~0ull; }

private:
    void note_first(uint32_t pi, uint64_t n);
    void record(uint64_t n, uint64_t p, const char* kind, const char* what);
    std::string gap_fields() const;
    std::vector<uint64_t> by_prime(const std::vector<uint64_t>& cls, const std::vector<uint64_t>* agg) const;
    void load_state();
    void save_state() const;
    std::string state_text() const;
    struct GapCampaign {
        GapTotals g;
        uint64_t evens = 0, walked = 0;
        std::vector<std::pair<uint64_t, uint64_t>> covered;
    };
    GapCampaign load_gaps() const;
    bool commit_gaps();
    bool gaps_cover(uint64_t a, uint64_t b) const;
    bool log_has_gaps_in(uint64_t a, uint64_t b) const;
    void append_log(const std::string& line) const;
    void reset_campaign();

    const std::vector<uint32_t>& cands_;
    const uint32_t tail_t_;
    size_t np_ = 0;
    std::vector<int32_t> pidx_;                // p >> 1 -> prime index, -1 none
    std::string state_path_, log_path_, rec_path_, live_path_;
    bool campaign_ = true;
    // campaign
    std::vector<uint64_t> cum_samp_, cum_tail_, cum_tail_agg_;   // [cr * np + i], [cr * np + i], [i]
    std::vector<uint64_t> first_;
    std::vector<std::pair<uint64_t, uint64_t>> covered_, tot_beyond_;
    uint64_t tot_evens_ = 0, tot_sampled_ = 0, tot_primes_[8] = {}, tot_twins_[3] = {};
    // run
    std::vector<uint64_t> samp_, tail_, tail_agg_;
    std::vector<uint64_t> run_first_;
    std::vector<std::pair<uint64_t, uint64_t>> beyond_;      // (n, p)
    uint64_t primes_res_[8] = {}, twins_[3] = {};
    GapTotals gaps_;
    uint64_t gap_last_ = 0, gap_start_ = ~0ull;
    bool gap_last_walked_ = false;
    uint64_t walk_evens_ = 0;
    uint64_t start_ = 0, limit_ = 0, evens_ = 0, evens_sampled_ = 0;
    std::vector<uint8_t> want_;
    std::vector<uint16_t> unknown_;
    std::vector<std::pair<uint64_t, uint64_t>> events_;
    long long live_t_ = 0;
};

}  // namespace gb
