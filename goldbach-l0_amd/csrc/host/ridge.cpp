// Goldbach ridge tracker and store (contracts in ridge.hpp). The store
// format, validation rules and ordering follow the reference's
// include/ridge.hpp:52-236 (an existing campaign's goldbach_ridges.csv keeps
// loading).
#include "ridge.hpp"

#include <algorithm>
#include <cstdio>
#include <ctime>
#include <mutex>
#include <unordered_map>

#include "env.hpp"
#include "primes.hpp"

namespace gb {
namespace {

// larger ridge first, equal ridge: larger even first
bool before(const RidgeEntry& a, const RidgeEntry& b) { return a.p != b.p ? a.p > b.p : a.n > b.n; }

// The exact checks' verdicts on (n, p) pairs already checked in this process
// (a store reloaded run after run re-checks the same rows: each costs up to
// a few hundred Miller-Rabin tests). The verdict depends only on n, p and the
// candidate list, which is the same fixed prime list for every tracker.
enum Verdict : uint8_t { kOk, kBadPair, kNotMinimal };
struct PairHash {
    size_t operator()(const std::pair<uint64_t, uint64_t>& k) const
    {
        return size_t(k.first * 0x9E3779B97F4A7C15ull ^ (k.second + (k.first >> 29)));
    }
};
std::mutex g_memo_mu;
std::unordered_map<std::pair<uint64_t, uint64_t>, uint32_t, PairHash> g_memo;   // verdict | witness << 8
constexpr size_t kMemoMax = 1 << 16;

}  // namespace

RidgeTracker::RidgeTracker(const std::vector<uint32_t>* cand) : cand_(cand)
{
    if (!env::onoff("GB_RIDGE", true)) return;
    floor_ = env::u32("GB_RIDGE_MIN", RIDGE_MIN_DEFAULT);
    if (floor_ < 3) floor_ = 3;
    debug_ = env::present("GB_RIDGE_DEBUG");
}

void RidgeTracker::clamp_floor_above(uint32_t bound)
{
    if (floor_ != 0 && floor_ <= bound) {
        floor_ = bound + 1;
        std::fprintf(stderr, "note: ridge floor raised to %u (phase-A candidates reach %u)\n", floor_, bound);
    }
}

uint32_t RidgeTracker::thresh() const
{
    if (!active()) return 0xFFFFFFFFu;
    return top_.size() >= RIDGE_KEEP ? std::max<uint32_t>(floor_, uint32_t(top_.back().p)) : floor_;
}

void RidgeTracker::consider(uint64_t n, uint64_t p, uint64_t seg, uint64_t found)
{
    if (!active() || p < floor_) return;
    ++appends_;
    if (n < 4 || (n & 1) || p < 3 || !(p & 1) || p >= n) {
        ++dropped_;
        return;
    }
    if (top_.size() >= RIDGE_KEEP) {          // cheap qualification before the exact checks
        const RidgeEntry& last = top_.back();
        if (p < last.p || (p == last.p && n <= last.n)) return;
    }
    for (const RidgeEntry& t : top_)
        if (t.n == n) return;                 // already stored: keep its found time
    const std::pair<uint64_t, uint64_t> key{n, p};
    uint32_t v = UINT32_MAX;
    if (cand_) {
        std::lock_guard<std::mutex> lk(g_memo_mu);
        const auto it = g_memo.find(key);
        if (it != g_memo.end()) v = it->second;
    }
    if (v == UINT32_MAX) {
        v = kOk;
        if (!is_prime_u64(p) || !is_prime_u64(n - p)) {
            v = kBadPair;
        } else if (cand_) {
            for (uint32_t c : *cand_) {       // minimality: no smaller candidate works
                if (c >= p) break;
                if (is_prime_u64(n - c)) {
                    v = kNotMinimal | c << 8;
                    break;
                }
            }
        }
        if (cand_) {
            std::lock_guard<std::mutex> lk(g_memo_mu);
            if (g_memo.size() >= kMemoMax) g_memo.clear();
            g_memo.emplace(key, v);
        }
    }
    if ((v & 0xFF) == kBadPair) {
        ++dropped_;
        if (debug_) std::fprintf(stderr, "ridge drop (pair): n=%llu p=%llu\n", (unsigned long long)n,
                                 (unsigned long long)p);
        return;
    }
    if ((v & 0xFF) == kNotMinimal) {
        ++dropped_;
        if (debug_) std::fprintf(stderr, "ridge drop (not minimal): n=%llu p=%llu < %u\n", (unsigned long long)n,
                                 (unsigned long long)p, v >> 8);
        return;
    }
    if (found == UINT64_MAX) found = uint64_t(std::time(nullptr));
    const RidgeEntry e{n, p, n - p, seg, found};
    top_.insert(std::upper_bound(top_.begin(), top_.end(), e, before), e);
    if (top_.size() > RIDGE_KEEP) top_.pop_back();
    changed_ = true;
}

void RidgeTracker::merge(const uint64_t* pairs, uint32_t cnt, uint32_t cap, uint64_t seg)
{
    if (!active()) return;
    if (cnt > cap) overflow_ = true;
    const uint32_t n = std::min(cnt, cap);
    for (uint32_t i = 0; i < n; ++i) consider(pairs[2 * size_t(i)], pairs[2 * size_t(i) + 1], seg);
}

bool RidgeTracker::load(const std::string& path)
{
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        unsigned long long r, n, p, q, s, t = 0;   // t is absent in v1 rows
        if (std::sscanf(line, "%llu,%llu,%llu,%llu,%llu,%llu", &r, &n, &p, &q, &s, &t) >= 5) consider(n, p, s, t);
    }
    std::fclose(f);
    changed_ = false;
    return true;
}

void RidgeTracker::save(const std::string& path)
{
    if (!changed_) return;
    std::string out = "# goldbach-l0 ridge store v2 - top 30 largest Goldbach ridges\n"
                      "# rank,even,small_prime_p,large_prime_q,segment_start,found_unix\n";
    char row[192];
    for (size_t i = 0; i < top_.size(); ++i) {
        const RidgeEntry& e = top_[i];
        std::snprintf(row, sizeof row, "%zu,%llu,%llu,%llu,%llu,%llu\n", i + 1, (unsigned long long)e.n,
                      (unsigned long long)e.p, (unsigned long long)e.q, (unsigned long long)e.seg,
                      (unsigned long long)e.found);
        out += row;
    }
    const std::string tmp = path + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "ridge store: cannot write %s\n", tmp.c_str());
        return;
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) == 0 && ok) std::rename(tmp.c_str(), path.c_str());
    else std::remove(tmp.c_str());
    changed_ = false;
}

}  // namespace gb
