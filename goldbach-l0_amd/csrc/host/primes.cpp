// Host prime utilities (contracts in primes.hpp).

#include "primes.hpp"

#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <filesystem>
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

// ---------------------------------------------------------------- prime cache
// GBP2 file: "GBP2", u64 limit, u64 count, u64 payload hash, then count u32
// odd primes ascending (the reference's format, so caches are shared).

uint64_t prime_payload_hash(const uint32_t* p, size_t cnt)
{
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < cnt; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

namespace {

constexpr char kCacheMagic[4] = {'G', 'B', 'P', '2'};
constexpr size_t kCacheHeader = 28;

// the odd primes below 64 first, then strictly ascending odd values <= lim
bool payload_ok(const uint32_t* p, size_t cnt, uint64_t lim)
{
    static constexpr uint32_t head[] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
    for (size_t i = 0; i < cnt && i < sizeof head / sizeof head[0]; ++i)
        if (p[i] != head[i]) return false;
    for (size_t i = 1; i < cnt; ++i)
        if (p[i] <= p[i - 1] || !(p[i] & 1u)) return false;
    return cnt == 0 || p[cnt - 1] <= lim;
}

// The primes <= n of the cache file, when its limit is n (or larger, with
// allow_larger) and it passes every check.
bool read_cache(const std::string& file, uint64_t n, bool allow_larger, std::vector<uint32_t>& out)
{
    std::FILE* f = std::fopen(file.c_str(), "rb");
    if (!f) return false;
    char magic[4] = {};
    uint64_t lim = 0, cnt = 0, hash = 0;
    const char* why = nullptr;
    if (std::fread(magic, 1, 4, f) != 4 || std::fread(&lim, 8, 1, f) != 1 || std::fread(&cnt, 8, 1, f) != 1 ||
        std::fread(&hash, 8, 1, f) != 1)
        why = "short header";
    else if (std::memcmp(magic, kCacheMagic, 4) != 0)
        why = "unknown format";
    else if (!(lim == n || (allow_larger && lim > n))) {
        std::fclose(f);                  // another limit: not this run's file
        return false;
    } else if (cnt == 0 || cnt >= (uint64_t(1) << 32))
        why = "size does not match the header";
    if (!why) {
        std::vector<uint32_t> v(cnt);
        if (std::fread(v.data(), 4, cnt, f) != cnt || std::fgetc(f) != EOF) why = "size does not match the header";
        else if (prime_payload_hash(v.data(), cnt) != hash) why = "payload hash mismatch";
        else if (!payload_ok(v.data(), cnt, lim)) why = "payload content check failed";
        if (!why) {
            std::fclose(f);
            if (lim > n) v.resize(size_t(std::upper_bound(v.begin(), v.end(), uint32_t(std::min<uint64_t>(n, UINT32_MAX))) -
                                         v.begin()));
            out.swap(v);
            return true;
        }
    }
    std::fclose(f);
    std::fprintf(stderr, "warning: prime cache %s ignored (%s); regenerating\n", file.c_str(), why);
    return false;
}

void write_cache(const std::string& path, uint64_t n, const std::vector<uint32_t>& primes)
{
    const std::string tmp = path + ".tmp." + std::to_string(::getpid());
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const uint64_t cnt = primes.size(), h = prime_payload_hash(primes.data(), primes.size());
    bool ok = std::fwrite(kCacheMagic, 1, 4, f) == 4 && std::fwrite(&n, 8, 1, f) == 1 &&
              std::fwrite(&cnt, 8, 1, f) == 1 && std::fwrite(&h, 8, 1, f) == 1 &&
              std::fwrite(primes.data(), 4, primes.size(), f) == primes.size();
    ok = std::fflush(f) == 0 && ok;
    ok = ::fsync(::fileno(f)) == 0 && ok;
    ok = std::fclose(f) == 0 && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) std::remove(tmp.c_str());
}

}  // namespace

std::vector<uint32_t> base_primes_cached(uint32_t n, const char* cache_path, unsigned threads)
{
    if (!cache_path || !*cache_path) return base_primes(n, threads);
    std::vector<uint32_t> v;
    const std::string keyed = std::string(cache_path) + "." + std::to_string(n);
    if (read_cache(keyed, n, false, v)) return v;
    {   // the smallest keyed cache above n: its prefix
        namespace fs = std::filesystem;
        const fs::path p(cache_path);
        const fs::path dir = p.has_parent_path() ? p.parent_path() : fs::path(".");
        const std::string stem = p.filename().string() + ".";
        uint64_t best = 0;
        std::error_code ec;
        for (const auto& de : fs::directory_iterator(dir, ec)) {
            const std::string name = de.path().filename().string();
            if (name.size() <= stem.size() || name.compare(0, stem.size(), stem) != 0) continue;
            const std::string num = name.substr(stem.size());
            if (num.find_first_not_of("0123456789") != std::string::npos) continue;
            const uint64_t L = std::strtoull(num.c_str(), nullptr, 10);
            if (L > n && (best == 0 || L < best)) best = L;
        }
        if (best && read_cache((dir / (stem + std::to_string(best))).string(), n, true, v)) return v;
    }
    v = base_primes(n, threads);
    write_cache(keyed, n, v);
    return v;
}

}  // namespace gb
