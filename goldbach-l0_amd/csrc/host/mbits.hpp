#pragma once
// The resident multiplier bitmap of the large-prime stage (the reference's
// multiplier table, built by mt_build at start-up: goldbach.cl:3181-3242,
// src/verifier.cpp:1444-1508). Here it covers every multiplier the run can
// need, up to a device-memory budget, instead of the reference's 3.5 GiB cap.
//
// Layout (gbh_mtrans_args): word w, bit 8 j + i <-> m = m_base + 240 w + 30 j
// + R[i], R the eight unit residues mod 30 ascending; a bit is set iff m is
// prime. One bitmap per process: engines share it, and a run that needs
// multipliers outside it rebuilds it over the union (within the budget).
#include <cstdint>
#include <memory>
#include <vector>

namespace gb {

struct MultBitmap {
    const uint64_t* bits = nullptr;   // device
    uint64_t m_base = 0;              // a multiple of 240
    uint64_t m_span = 0;              // covers m in [m_base, m_base + m_span)
    uint64_t words = 0;
    double build_s = 0;               // seconds the build took
    bool covers(uint64_t lo, uint64_t hi) const { return bits && lo >= m_base && hi < m_base + m_span; }
    ~MultBitmap();
    void* owner = nullptr;            // the allocating context
};

/// Device-memory budget of the bitmap in bytes (GB_MT_GB, GiB; default 80;
/// 0 turns the bitmap off and every large prime walks the wheel).
uint64_t mult_bitmap_budget();

/// A bitmap covering [m_lo, m_hi] (m_lo > 61), built with the odd primes of
/// @p base (ascending, from 3, reaching sqrt(m_hi)) when the cached one does
/// not cover the range. @throws std::runtime_error when the range exceeds
/// the budget or the base primes stop below its square root.
std::shared_ptr<const MultBitmap> mult_bitmap(uint64_t m_lo, uint64_t m_hi, const std::vector<uint32_t>& base,
                                              bool verbose);

/// The cached bitmap, if any (no build).
std::shared_ptr<const MultBitmap> mult_bitmap_peek();

/// Drops the process's cached bitmap (freed once no run holds it).
void mult_bitmap_release();

}  // namespace gb
