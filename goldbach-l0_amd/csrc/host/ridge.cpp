// Goldbach ridge board and store (contracts and format in ridge.hpp).
#include "ridge.hpp"

#include <algorithm>
#include <cstdio>
#include <ctime>
#include <mutex>
#include <string_view>
#include <unordered_map>

#include "env.hpp"
#include "primes.hpp"
#include "textio.hpp"

namespace gb {
namespace {

// ordering of the board: the larger ridge first; equal ridges by the
// larger even first
bool ranks_above(const RidgeEntry& a, const RidgeEntry& b) { return a.p > b.p || (a.p == b.p && a.n > b.n); }

// Outcome of the exact check of a pair.
struct Check {
    enum Kind : uint8_t { Ok, NotPrime, Smaller } kind = Ok;
    uint32_t smaller = 0;   // a smaller candidate with n - p' prime (Smaller)
};

// The exact check: n even, p and n - p prime, and p the smallest candidate
// with n - p prime. A store reloaded run after run checks the same rows
// again (each up to a few hundred Miller-Rabin tests), so the outcomes are
// remembered for the process: they depend only on (n, p) and the candidate
// list, the same fixed list for every board.
class CheckMemo {
public:
    Check get(uint64_t n, uint64_t p, const std::vector<uint32_t>* cand)
    {
        const Key k{n, p};
        if (cand) {
            std::lock_guard<std::mutex> g(mu_);
            if (auto it = seen_.find(k); it != seen_.end()) return it->second;
        }
        Check c;
        if (!is_prime_u64(p) || !is_prime_u64(n - p)) {
            c.kind = Check::NotPrime;
        } else if (cand) {
            const auto end = std::lower_bound(cand->begin(), cand->end(), p);
            const auto hit = std::find_if(cand->begin(), end, [n](uint32_t q) { return is_prime_u64(n - q); });
            if (hit != end) c = Check{Check::Smaller, *hit};
        }
        if (cand) {
            std::lock_guard<std::mutex> g(mu_);
            if (seen_.size() >= kCap) seen_.clear();
            seen_.emplace(k, c);
        }
        return c;
    }

private:
    struct Key {
        uint64_t n, p;
        bool operator==(const Key& o) const { return n == o.n && p == o.p; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const { return size_t((k.n ^ (k.p << 40) ^ (k.p >> 24)) * 0xD6E8FEB86659FD93ull); }
    };
    static constexpr size_t kCap = 1 << 16;
    std::mutex mu_;
    std::unordered_map<Key, Check, KeyHash> seen_;
};

CheckMemo& memo()
{
    static CheckMemo m;
    return m;
}

}  // namespace

RidgeTracker::RidgeTracker(const std::vector<uint32_t>* cand) : cand_(cand)
{
    if (!env::onoff("GB_RIDGE", true)) return;
    floor_ = std::max<uint32_t>(env::u32("GB_RIDGE_MIN", RIDGE_MIN_DEFAULT), 3);
    trace_ = env::present("GB_RIDGE_DEBUG");
}

void RidgeTracker::clamp_floor_above(uint32_t bound)
{
    if (!active() || floor_ > bound) return;
    floor_ = bound + 1;
    std::fprintf(stderr, "note: ridge floor raised to %u (phase-A candidates reach %u)\n", floor_, bound);
}

uint32_t RidgeTracker::thresh() const
{
    if (!active()) return 0xFFFFFFFFu;
    if (board_.size() < RIDGE_KEEP) return floor_;
    return std::max<uint32_t>(floor_, uint32_t(board_.back().p));
}

bool RidgeTracker::admit(const RidgeEntry& e)
{
    // a full board takes only a ridge that ranks above its last one
    if (board_.size() >= RIDGE_KEEP && !ranks_above(e, board_.back())) return false;
    if (std::any_of(board_.begin(), board_.end(), [&](const RidgeEntry& b) { return b.n == e.n; }))
        return false;   // already on the board: its found time stays
    const Check c = memo().get(e.n, e.p, cand_);
    if (c.kind != Check::Ok) {
        ++dropped_;
        if (trace_)
            std::fprintf(stderr, "ridge drop (%s): n=%llu p=%llu%s\n", c.kind == Check::NotPrime ? "pair" : "not minimal",
                         (unsigned long long)e.n, (unsigned long long)e.p,
                         c.kind == Check::Smaller ? (" > " + std::to_string(c.smaller)).c_str() : "");
        return false;
    }
    board_.insert(std::upper_bound(board_.begin(), board_.end(), e, ranks_above), e);
    if (board_.size() > RIDGE_KEEP) board_.resize(RIDGE_KEEP);
    dirty_ = true;
    return true;
}

void RidgeTracker::consider(uint64_t n, uint64_t p, uint64_t seg, uint64_t found)
{
    if (!active() || p < floor_) return;
    ++appends_;
    const bool shape_ok = n >= 4 && n % 2 == 0 && p >= 3 && p % 2 == 1 && p < n;
    if (!shape_ok) {
        ++dropped_;
        return;
    }
    (void)admit(RidgeEntry{n, p, n - p, seg, found == UINT64_MAX ? uint64_t(std::time(nullptr)) : found});
}

void RidgeTracker::merge(const uint64_t* pairs, size_t stored, uint64_t claimed, uint64_t seg)
{
    if (!active()) return;
    overflowed_ = overflowed_ || claimed > stored;
    const size_t n = size_t(std::min<uint64_t>(claimed, stored));
    for (size_t i = 0; i < n; ++i) consider(pairs[2 * i], pairs[2 * i + 1], seg);
}

bool RidgeTracker::load(const std::string& path)
{
    const std::optional<std::string> text = textio::read_file(path);
    if (!text) return false;
    std::string_view rest(*text);
    while (!rest.empty()) {
        const size_t nl = rest.find('\n');
        std::string_view line = rest.substr(0, nl);
        rest = nl == std::string_view::npos ? std::string_view() : rest.substr(nl + 1);
        if (!line.empty() && line.back() == '\r') line.remove_suffix(1);
        if (line.empty() || line.front() == '#') continue;
        // rank,n,p,q,segment_start[,found_unix]: the rank and q are
        // recomputed, the rest is checked again before it counts
        uint64_t v[6] = {};
        size_t cols = 0;
        bool ok = true;
        while (ok && cols < 6) {
            const size_t comma = line.find(',');
            ok = textio::parse(line.substr(0, comma), v[cols]);
            cols += ok;
            if (comma == std::string_view::npos) break;
            line.remove_prefix(comma + 1);
        }
        if (cols >= 5) consider(v[1], v[2], v[4], cols == 6 ? v[5] : 0);
    }
    dirty_ = false;
    return true;
}

void RidgeTracker::save(const std::string& path)
{
    if (!dirty_) return;
    std::string out = "# goldbach-l0 ridge store v2 - top 30 largest Goldbach ridges\n"
                      "# rank,even,small_prime_p,large_prime_q,segment_start,found_unix\n";
    for (size_t i = 0; i < board_.size(); ++i) {
        const RidgeEntry& e = board_[i];
        out += std::to_string(i + 1) + "," + std::to_string(e.n) + "," + std::to_string(e.p) + "," +
               std::to_string(e.q) + "," + std::to_string(e.seg) + "," + std::to_string(e.found) + "\n";
    }
    try {
        textio::replace_file(path, out);
        dirty_ = false;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "ridge store: %s\n", ex.what());
    }
}

}  // namespace gb
