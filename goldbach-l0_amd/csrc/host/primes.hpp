#pragma once
// Host prime utilities of the verifier: base primes for the sieve, the
// deterministic 64-bit Miller-Rabin and the CPU witness search of the
// leftover fallback (reference include/primes.hpp:180-255, src/primes.cpp).
#include <cmath>
#include <functional>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gb {

/// Largest limit accepted: 2^64 - 2^40 (primes.hpp:212-222): first-hit
/// arithmetic adds less than 2^40 to values <= limit.
inline constexpr uint64_t MAX_SUPPORTED_LIMIT = 0ull - (1ull << 40);

inline void require_supported_limit(uint64_t limit)
{
    if (limit > MAX_SUPPORTED_LIMIT)
        throw std::runtime_error("limit " + std::to_string(limit) + " is above the supported maximum 2^64 - 2^40 = " +
                                 std::to_string(MAX_SUPPORTED_LIMIT));
}

/// floor(sqrt(limit)) + 16, saturated at 2^32 - 1 (primes.hpp:224-235).
inline uint32_t base_prime_bound(uint64_t limit)
{
    const double r = std::sqrt(double(limit)) + 16.0;
    return r >= 4294967295.0 ? 0xFFFFFFFFu : uint32_t(r);
}

/// The odd primes 3 <= p <= n, ascending.
std::vector<uint32_t> odd_primes_up_to(uint32_t n);

/// The odd primes 3 <= p <= n by a segmented, multi-threaded sieve (the
/// base primes: up to ~2^32, ~200M entries).
std::vector<uint32_t> base_primes(uint32_t n, unsigned threads);

/// FNV-1a (64-bit) over the 32-bit words of a prime payload: the content
/// hash of the prime cache (primes.cpp:113-122 of the reference).
uint64_t prime_payload_hash(const uint32_t* p, size_t cnt);

/// base_primes through the reference's prime cache (GBP2 files,
/// src/primes.cpp:125-284): <path>.<n> when it exists and passes its checks,
/// else the smallest <path>.<L> with L > n (its primes <= n), else the sieve,
/// published as <path>.<n> (temp file, fsync, rename; best effort). A cache
/// that fails its header, hash or content check is reported and ignored.
std::vector<uint32_t> base_primes_cached(uint32_t n, const char* cache_path, unsigned threads);

/// The large primes as the harvest's packed stream (gbh_pstream: groups of
/// @p group primes, each group's first prime as its base, one byte of half
/// gap for every other prime; the reference's pack_large_primes,
/// src/host_math.cpp:186-209), padded by @p pad entries past the last prime
/// (gaps 0, bases the last prime). Throws when a half gap exceeds 255.
struct PackedPrimes {
    std::vector<uint32_t> base;
    std::vector<uint8_t> gap;
    uint32_t n = 0;
};
PackedPrimes pack_primes(const uint32_t* p, uint32_t n, uint32_t group, uint32_t pad);

/// Prime j of a packed stream (host mirror of the device decode).
uint32_t packed_prime_at(const PackedPrimes& pk, uint32_t j, uint32_t group);

/// Deterministic Miller-Rabin (bases 2..37): exact for every uint64_t.
bool is_prime_u64(uint64_t n);

/// Smallest odd prime p with n - p prime: the candidates @p cands first
/// (ascending odd primes up to @p cand_bound), then every odd p above it up
/// to n / 2 (primes.cpp:482-493). 2 for n = 4; 0 when none exists.
uint64_t cpu_min_goldbach_prime(uint64_t n, const uint32_t* cands, size_t ncand, uint64_t cand_bound);

/// Largest prime < x (x > 3).
uint64_t prev_prime(uint64_t x);

/// Runs fn(i) for i in [0, n) on up to @p threads threads (fn must not throw).
void parallel_for(uint64_t n, unsigned threads, const std::function<void(uint64_t)>& fn);

}  // namespace gb
