#pragma once
// Goldbach ridges: the largest smallest-partition primes p(n) a campaign has
// met, kept as a persistent top-30 store. The store's CSV format is the
// reference's ridge store v2 (include/ridge.hpp:203-225: a header, then
// rank,even,p,q,segment_start,found_unix rows; v1 rows lack found_unix), so
// an existing campaign's goldbach_ridges.csv keeps loading.
//
// A ridge (n, p) enters the board only after an exact check: p and n - p
// prime (deterministic Miller-Rabin) and no smaller candidate p' with
// n - p' prime. Phase B records every pending even whose p(n) reaches the
// device threshold (the board's 30th ridge, at least the floor); every
// phase-A candidate lies below the floor, so above it the records are
// complete. GB_RIDGE=0 disables tracking, GB_RIDGE_MIN sets the floor
// (default 3000, raised above the phase-A candidates), GB_RIDGE_FILE the
// store (default goldbach_ridges.csv).
#include <cstdint>
#include <string>
#include <vector>

#include "config.hpp"
#include "verifier.hpp"

namespace gb {

class RidgeTracker {
public:
    /// @param cand ascending odd candidate primes (for the minimality check)
    explicit RidgeTracker(const std::vector<uint32_t>* cand);
    bool active() const { return floor_ != 0; }
    uint32_t floor() const { return floor_; }
    /// Raises the floor above @p bound (the largest phase-A candidate).
    void clamp_floor_above(uint32_t bound);
    /// The device's recording threshold: the board's 30th ridge when it is
    /// full (ties still recorded), at least the floor.
    uint32_t thresh() const;
    /// One witness (n, p) found in the segment starting at @p seg; checked
    /// exactly before it may enter the board. @p found: discovery time
    /// (Unix seconds), UINT64_MAX = now.
    void consider(uint64_t n, uint64_t p, uint64_t seg, uint64_t found = UINT64_MAX);
    /// A segment's device records: @p claimed recorded by the device, the
    /// first @p stored of them in @p pairs as (n, p); claimed > stored marks
    /// the run as overflowed (the excess was not considered).
    void merge(const uint64_t* pairs, size_t stored, uint64_t claimed, uint64_t seg);
    /// Loads the store, every row checked again; false when unreadable.
    bool load(const std::string& path);
    /// Rewrites the store (atomically, synced) when the board changed.
    void save(const std::string& path);
    uint64_t appends() const { return appends_; }
    uint64_t dropped() const { return dropped_; }
    bool overflowed() const { return overflowed_; }
    const std::vector<RidgeEntry>& top() const { return board_; }

private:
    bool admit(const RidgeEntry& e);

    const std::vector<uint32_t>* cand_;
    std::vector<RidgeEntry> board_;   // best first: larger p, then larger n
    uint32_t floor_ = 0;
    bool trace_ = false;
    uint64_t appends_ = 0, dropped_ = 0;
    bool overflowed_ = false, dirty_ = false;
};

}  // namespace gb
