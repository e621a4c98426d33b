// Base primes through a GBP2 prime cache (contract in primes.hpp).
//
// A GBP2 file holds the odd primes up to a bound, shared with the
// reference's own caches (its format, src/primes.cpp:125-284 of the
// reference): the 4 bytes "GBP2", then little-endian u64 bound, u64 count
// and u64 FNV-1a hash of the payload's 32-bit words, then count u32 primes
// ascending. A run needing the primes up to n looks for <stem>.<n>, then for
// the smallest <stem>.<L> with L > n (whose prefix serves it), else sieves
// and publishes <stem>.<n>. A file that fails any check is reported and
// passed over.
#include <dirent.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <optional>
#include <string>

#include "primes.hpp"

namespace gb {

uint64_t prime_payload_hash(const uint32_t* p, size_t cnt)
{
    constexpr uint64_t kOffset = 14695981039346656037ull, kPrime = 1099511628211ull;
    uint64_t h = kOffset;
    for (const uint32_t* e = p + cnt; p != e; ++p) h = (h ^ *p) * kPrime;
    return h;
}

namespace {

#pragma pack(push, 1)
struct Gbp2Header {
    char magic[4];
    uint64_t bound, count, hash;
};
#pragma pack(pop)
static_assert(sizeof(Gbp2Header) == 28, "GBP2 header is 28 bytes");

bool read_exact(int fd, void* dst, size_t bytes)
{
    auto* p = static_cast<char*>(dst);
    while (bytes) {
        const ssize_t r = ::read(fd, p, bytes);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        bytes -= size_t(r);
    }
    return true;
}

bool write_exact(int fd, const void* src, size_t bytes)
{
    auto* p = static_cast<const char*>(src);
    while (bytes) {
        const ssize_t w = ::write(fd, p, bytes);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) return false;
        p += w;
        bytes -= size_t(w);
    }
    return true;
}

// The payload is a plausible prime list: it opens with the odd primes below
// 64, then strictly ascending odd values up to the bound.
const char* payload_fault(const std::vector<uint32_t>& v, uint64_t bound)
{
    static constexpr uint32_t kOpening[] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};
    const size_t k = std::min(v.size(), sizeof kOpening / sizeof kOpening[0]);
    if (!std::equal(v.begin(), v.begin() + k, kOpening)) return "payload content check failed";
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i] <= v[i - 1] || (v[i] & 1u) == 0) return "payload content check failed";
    if (!v.empty() && v.back() > bound) return "payload content check failed";
    return nullptr;
}

enum class Want { Exact, AtLeast };

// The primes <= n from one cache file. nullopt when the file is absent, is
// for another bound, or fails a check (then reported).
std::optional<std::vector<uint32_t>> load_cache(const std::string& file, uint32_t n, Want want)
{
    const int fd = ::open(file.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) return std::nullopt;
    struct stat st{};
    Gbp2Header h{};
    const char* fault = nullptr;
    std::vector<uint32_t> v;
    if (::fstat(fd, &st) != 0 || !read_exact(fd, &h, sizeof h)) {
        fault = "short header";
    } else if (std::memcmp(h.magic, "GBP2", 4) != 0) {
        fault = "unknown format";
    } else if (want == Want::Exact ? h.bound != n : h.bound <= n) {
        ::close(fd);   // a cache for another bound: not this run's file
        return std::nullopt;
    } else if (h.count == 0 || h.count >= (uint64_t(1) << 32) || uint64_t(st.st_size) != sizeof h + 4 * h.count) {
        fault = "size does not match the header";
    } else {
        v.resize(size_t(h.count));
        if (!read_exact(fd, v.data(), v.size() * 4)) fault = "size does not match the header";
        else if (prime_payload_hash(v.data(), v.size()) != h.hash) fault = "payload hash mismatch";
        else fault = payload_fault(v, h.bound);
    }
    ::close(fd);
    if (fault) {
        std::fprintf(stderr, "warning: prime cache %s ignored (%s); regenerating\n", file.c_str(), fault);
        return std::nullopt;
    }
    v.erase(std::upper_bound(v.begin(), v.end(), n), v.end());   // a larger bound's prefix
    return v;
}

// The smallest bound L > n among the files <stem>.<L> of the stem's
// directory (0: none).
uint64_t next_larger_bound(const std::string& stem, uint32_t n)
{
    const size_t slash = stem.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : stem.substr(0, slash + 1);
    const std::string prefix = (slash == std::string::npos ? stem : stem.substr(slash + 1)) + ".";
    DIR* d = ::opendir(dir.c_str());
    if (!d) return 0;
    uint64_t best = 0;
    while (const dirent* e = ::readdir(d)) {
        const std::string name = e->d_name;
        if (name.size() <= prefix.size() || name.compare(0, prefix.size(), prefix) != 0) continue;
        const std::string digits = name.substr(prefix.size());
        if (digits.size() > 19 || digits.find_first_not_of("0123456789") != std::string::npos) continue;
        const uint64_t L = std::stoull(digits);
        if (L > n && (best == 0 || L < best)) best = L;
    }
    ::closedir(d);
    return best;
}

// Publishes the primes as <file> (bound n): a private temp file, synced,
// renamed into place. Best effort: a cache that cannot be written only
// costs the next run a sieve.
void publish_cache(const std::string& file, uint32_t n, const std::vector<uint32_t>& v)
{
    const std::string part = file + ".part-" + std::to_string(::getpid());
    const int fd = ::open(part.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (fd < 0) return;
    Gbp2Header h{};
    std::memcpy(h.magic, "GBP2", 4);
    h.bound = n;
    h.count = v.size();
    h.hash = prime_payload_hash(v.data(), v.size());
    const bool ok = write_exact(fd, &h, sizeof h) && write_exact(fd, v.data(), v.size() * 4) && ::fsync(fd) == 0;
    if (::close(fd) != 0 || !ok || ::rename(part.c_str(), file.c_str()) != 0) ::unlink(part.c_str());
}

}  // namespace

std::vector<uint32_t> base_primes_cached(uint32_t n, const char* cache_path, unsigned threads)
{
    if (!cache_path || !*cache_path) return base_primes(n, threads);
    const std::string stem = cache_path, own = stem + "." + std::to_string(n);
    if (auto v = load_cache(own, n, Want::Exact)) return std::move(*v);
    if (const uint64_t L = next_larger_bound(stem, n))
        if (auto v = load_cache(stem + "." + std::to_string(L), n, Want::AtLeast)) return std::move(*v);
    std::vector<uint32_t> v = base_primes(n, threads);
    publish_cache(own, n, v);
    return v;
}

}  // namespace gb
