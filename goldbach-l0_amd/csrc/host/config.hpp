#pragma once
// Constants of the verifier surface (the reference's include/config.hpp:18-164
// values that shape results or the output format) and this build's own
// MI355X tuning defaults.
#include <cstdint>

namespace gb {

inline constexpr const char* VERSION = "2.8.0-mi355x";

// ---- kept from the reference (results and formats depend on them) --------
inline constexpr uint32_t MAX_GB_PRIME = 9781;                  // config.hpp:86
inline constexpr uint64_t DEFAULT_LIMIT = 100'000'000ULL;       // config.hpp:155
inline constexpr uint64_t DEFAULT_SEGMENT_EVEN = 467'500'000ULL;  // config.hpp:163 (segment span; sampling follows it)
inline constexpr uint32_t MAX_LEFTOVERS = 1u << 20;             // config.hpp:164
inline constexpr uint32_t RIDGE_CAP = 65536;                    // config.hpp:103
inline constexpr uint32_t RIDGE_KEEP = 30;                      // config.hpp:104
inline constexpr uint32_t RIDGE_MIN_DEFAULT = 3000;             // config.hpp:105
inline constexpr uint32_t CHECK_TAIL = 3000;                    // check30.hpp:42 (tail bound of the statistics)

// ---- MI355X pipeline defaults (this build) -------------------------------
// LDS window of the window sieve: 2^19 bits = 64 KB, two work-groups per CU.
inline constexpr uint32_t WIN_KB_DEFAULT = 64;
// Window primes are [67, WIN_P_DEFAULT); larger primes take the large stage.
inline constexpr uint64_t WIN_P_DEFAULT = uint64_t(1) << 19;   // window primes below this bound
// Runs of at least kLongBatch segments per batch also sieve the large wheel
// primes below WIN_P_LONG_DEFAULT in the windows.
inline constexpr uint64_t WIN_P_LONG_DEFAULT = uint64_t(1) << 20;
inline constexpr uint64_t kLongBatch = 8;
// Phase-A candidates per even class (the rest go to phase B).
inline constexpr uint32_t KA_DEFAULT = 96;

}  // namespace gb
