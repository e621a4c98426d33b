#pragma once
// Goldbach ridges: the largest smallest-partition primes p(n) a campaign has
// met, with a persistent top-30 store (the reference's RidgeTracker and
// store format, include/ridge.hpp:1-237).
//
// A ridge record (n, p) is kept only after exact validation: p and n - p
// prime (deterministic Miller-Rabin) and no smaller candidate p' with
// n - p' prime. Phase B records every pending even whose p(n) reaches the
// threshold (the current 30th largest ridge, at least the floor); every
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
    /// The device's recording threshold: the current 30th largest ridge
    /// (ties still recorded), at least the floor.
    uint32_t thresh() const;
    /// One witness record; validated exactly before it may enter the top.
    /// @p found: discovery time (Unix s), UINT64_MAX = now.
    void consider(uint64_t n, uint64_t p, uint64_t seg, uint64_t found = UINT64_MAX);
    /// A segment's device records: @p cnt claimed, the first min(cnt, cap)
    /// in @p pairs as (n, p).
    void merge(const uint64_t* pairs, uint32_t cnt, uint32_t cap, uint64_t seg);
    /// Loads the store (every row re-validated); false when unreadable.
    bool load(const std::string& path);
    /// Rewrites the store atomically when the top changed.
    void save(const std::string& path);
    uint64_t appends() const { return appends_; }
    uint64_t dropped() const { return dropped_; }
    bool overflowed() const { return overflow_; }
    const std::vector<RidgeEntry>& top() const { return top_; }

private:
    const std::vector<uint32_t>* cand_;
    std::vector<RidgeEntry> top_;
    uint32_t floor_ = 0;
    bool debug_ = false;
    uint64_t appends_ = 0, dropped_ = 0;
    bool overflow_ = false, changed_ = false;
};

}  // namespace gb
