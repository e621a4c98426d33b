// Host prime utilities (contracts in primes.hpp).

#include "primes.hpp"

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <thread>

namespace gb {

std::vector<uint32_t> odd_primes_up_to(uint32_t n)
{
    std::vector<uint32_t> out;
    if (n < 3) return out;
    std::vector<uint8_t> comp(n / 2 + 1, 0);            // index i <-> 2i + 1
    for (uint64_t p = 3; p * p <= n; p += 2)
        if (!comp[p / 2])
            for (uint64_t m = p * p; m <= n; m += 2 * p) comp[m / 2] = 1;
    for (uint64_t i = 1; 2 * i + 1 <= n; ++i)
        if (!comp[i]) out.push_back(uint32_t(2 * i + 1));
    return out;
}

void parallel_for(uint64_t n, unsigned threads, const std::function<void(uint64_t)>& fn)
{
    if (threads <= 1 || n <= 1) {
        for (uint64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> th;
    const unsigned nt = unsigned(std::min<uint64_t>(threads, n));
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (uint64_t i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto& x : th) x.join();
}

PackedPrimes pack_primes(const uint32_t* p, uint32_t n, uint32_t group, uint32_t pad)
{
    PackedPrimes pk;
    pk.n = n;
    if (!group || (group & (group - 1))) throw std::runtime_error("pack_primes: group not a power of two");
    const uint64_t groups = (uint64_t(n) + pad + group - 1) / group;
    if (groups * group > UINT32_MAX) throw std::runtime_error("pack_primes: stream beyond 2^32 entries");
    pk.base.assign(groups, n ? p[n - 1] : 0u);
    pk.gap.assign(groups * group, 0);
    for (uint32_t j = 0; j < n; ++j) {
        if (j % group == 0) {
            pk.base[j / group] = p[j];
            continue;
        }
        const uint32_t d = p[j] - p[j - 1];
        if (p[j] <= p[j - 1] || (d & 1) || d / 2 > 255)
            throw std::runtime_error("pack_primes: a gap of " + std::to_string(d) + " below " + std::to_string(p[j]) +
                                     " does not fit one byte of half gap");
        pk.gap[j] = uint8_t(d / 2);
    }
    return pk;
}

uint32_t packed_prime_at(const PackedPrimes& pk, uint32_t j, uint32_t group)
{
    const uint32_t g = j / group;
    uint32_t v = pk.base.at(g);
    for (uint32_t i = g * group + 1; i <= j; ++i) v += 2u * pk.gap.at(i);
    return v;
}

std::vector<uint32_t> base_primes(uint32_t n, unsigned threads)
{
    if (n < 1u << 20) return odd_primes_up_to(n);
    const uint32_t r = uint32_t(std::sqrt(double(n))) + 2;
    const std::vector<uint32_t> small = odd_primes_up_to(r);
    // blocks of 2^22 odd numbers: [lo, lo + 2^23)
    constexpr uint64_t kSpan = 1ull << 23;
    const uint64_t nblocks = (uint64_t(n) + kSpan) / kSpan;
    std::vector<std::vector<uint32_t>> parts(nblocks);
    parallel_for(nblocks, threads, [&](uint64_t b) {
        const uint64_t lo = b * kSpan, hi = std::min<uint64_t>(lo + kSpan, uint64_t(n) + 1);   // [lo, hi)
        std::vector<uint8_t> comp(kSpan / 2, 0);        // index i <-> lo + 2i + 1
        for (uint32_t p : small) {
            const uint64_t pp = uint64_t(p) * p;
            if (pp >= hi) break;
            uint64_t m = std::max<uint64_t>(pp, (lo + p - 1) / p * p);
            if (!(m & 1)) m += p;
            for (; m < hi; m += 2ull * p) comp[(m - lo) / 2] = 1;
        }
        auto& out = parts[b];
        for (uint64_t i = 0; lo + 2 * i + 1 < hi; ++i) {
            const uint64_t v = lo + 2 * i + 1;
            if (v >= 3 && !comp[i]) out.push_back(uint32_t(v));
        }
    });
    size_t total = 0;
    for (auto& p : parts) total += p.size();
    std::vector<uint32_t> out;
    out.reserve(total);
    for (auto& p : parts) out.insert(out.end(), p.begin(), p.end());
    return out;
}

namespace {
using u128 = unsigned __int128;

inline uint64_t mont_mul(uint64_t a, uint64_t b, uint64_t n, uint64_t ninv)
{
    const u128 t = u128(a) * b;
    const uint64_t m = uint64_t(t) * ninv;
    const u128 u = t + u128(m) * n;
    uint64_t hi = uint64_t(u >> 64);
    if (u < t || hi >= n) hi -= n;
    return hi;
}
}  // namespace

bool is_prime_u64(uint64_t n)
{
    static constexpr uint64_t kBases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t p : kBases)
        if (n % p == 0) return n == p;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    uint64_t x = n;                                     // n^-1 mod 2^64 by Newton steps
    for (int i = 0; i < 5; ++i) x *= 2 - n * x;
    const uint64_t ninv = 0 - x;
    const uint64_t one = uint64_t((u128(1) << 64) % n);
    const uint64_t r2 = uint64_t(u128(one) * one % n);
    const uint64_t mone = n - one;
    for (uint64_t a : kBases) {
        uint64_t b = mont_mul(a, r2, n, ninv), y = one;
        for (uint64_t e = d; e; e >>= 1) {
            if (e & 1) y = mont_mul(y, b, n, ninv);
            b = mont_mul(b, b, n, ninv);
        }
        if (y == one || y == mone) continue;
        bool comp = true;
        for (int i = 1; i < s && comp; ++i) {
            y = mont_mul(y, y, n, ninv);
            comp = y != mone;
        }
        if (comp) return false;
    }
    return true;
}

uint64_t cpu_min_goldbach_prime(uint64_t n, const uint32_t* cands, size_t ncand, uint64_t cand_bound)
{
    if (n < 4 || (n & 1)) return 0;
    if (n == 4) return 2;
    for (size_t i = 0; i < ncand; ++i) {
        const uint64_t p = cands[i];
        if (p > n - p) return 0;
        if (is_prime_u64(n - p)) return p;
    }
    for (uint64_t p = (cand_bound + 1) | 1; p <= n / 2; p += 2)
        if (is_prime_u64(p) && is_prime_u64(n - p)) return p;
    return 0;
}

uint64_t prev_prime(uint64_t x)
{
    uint64_t q = (x - 1) | 1;
    if (q >= x) q -= 2;
    while (!is_prime_u64(q)) q -= 2;
    return q;
}

}  // namespace gb
