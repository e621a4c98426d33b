// Exact statistics and campaign files (contracts and field semantics in
// gstats.hpp; formats as the reference's include/gstats.hpp).
#include "gstats.hpp"

#include <fcntl.h>
#include <sys/file.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <numeric>
#include <stdexcept>

#include "env.hpp"
#include "primes.hpp"

namespace gb {

std::string u128_str(u128 v)
{
    char b[48];
    int n = 0;
    do { b[n++] = char('0' + int(v % 10)); v /= 10; } while (v);
    std::string s;
    while (n) s += b[--n];
    return s;
}

namespace {

// Primes p <= 9781 with no even n <= 4e18 having p(n) = p (T. Oliveira e
// Silva's published S(p) table, the '?' entries; the reference lists them at
// include/gstats.hpp:127-136). A run's first even with such a p is reported
// as an S(p) candidate.
constexpr uint16_t kUnknownS[] = {
    8423, 8429, 8537, 8627, 8669, 8713, 8741, 8753, 8783, 8803, 8807, 8837, 8839, 8849, 8861,
    8863, 8867, 8887, 8893, 8923, 8963, 8969, 8999, 9007, 9011, 9013, 9029, 9041, 9043, 9049,
    9059, 9067, 9103, 9109, 9127, 9133, 9137, 9151, 9157, 9173, 9181, 9187, 9199, 9209, 9221,
    9227, 9239, 9241, 9257, 9277, 9281, 9283, 9293, 9311, 9319, 9323, 9337, 9343, 9349, 9371,
    9377, 9391, 9397, 9403, 9413, 9419, 9421, 9431, 9433, 9437, 9439, 9461, 9463, 9467, 9473,
    9479, 9491, 9497, 9511, 9521, 9533, 9539, 9547, 9551, 9587, 9601, 9613, 9619, 9623, 9631,
    9643, 9649, 9661, 9677, 9679, 9689, 9697, 9719, 9721, 9733, 9739, 9743, 9749, 9767, 9769};

int res_idx(uint64_t q)
{
    static const int idx[30] = {-1, 0, -1, -1, -1, -1, -1, 1, -1, -1, -1, 2, -1, 3, -1,
                                -1, -1, 4, -1, 5, -1, -1, -1, 6, -1, -1, -1, -1, -1, 7};
    const int i = idx[q % 30];
    if (i < 0) throw std::runtime_error("stats: prime " + std::to_string(q) + " outside the classes");
    return i;
}

int twin_idx(uint64_t q)
{
    const uint32_t m = uint32_t(q % 30);
    return m == 11 ? 0 : m == 17 ? 1 : 2;
}

void merge_ranges(std::vector<std::pair<uint64_t, uint64_t>>& r)
{
    std::sort(r.begin(), r.end());
    std::vector<std::pair<uint64_t, uint64_t>> m;
    for (auto& c : r) {
        if (!m.empty() && c.first <= m.back().second + 2) m.back().second = std::max(m.back().second, c.second);
        else m.push_back(c);
    }
    r = m;
}

// comma list by index, zero entries empty, trailing zeros cut
std::string dense_list(const std::vector<uint64_t>& v, size_t from = 0)
{
    size_t last = from;
    for (size_t i = from; i < v.size(); ++i)
        if (v[i]) last = i + 1;
    std::string s;
    for (size_t i = from; i < last; ++i) {
        if (i > from) s += ',';
        if (v[i]) s += std::to_string(v[i]);
    }
    return s;
}

std::vector<std::string> split_spaces(const std::string& line)
{
    std::vector<std::string> f;
    size_t pos = 0;
    for (;;) {
        const size_t e = line.find(' ', pos);
        f.push_back(line.substr(pos, e == std::string::npos ? std::string::npos : e - pos));
        if (e == std::string::npos) return f;
        pos = e + 1;
    }
}

bool parse_u128(const std::string& t, u128& v)
{
    if (t.empty()) return false;
    v = 0;
    const u128 mx = ~u128(0);
    for (char ch : t) {
        if (ch < '0' || ch > '9') return false;
        const unsigned d = unsigned(ch - '0');
        if (v > (mx - d) / 10) return false;
        v = v * 10 + d;
    }
    return true;
}
bool parse_u64(const std::string& t, uint64_t& v)
{
    u128 w = 0;
    if (!parse_u128(t, w) || w > u128(UINT64_MAX)) return false;
    v = uint64_t(w);
    return true;
}
bool parse_u32(const std::string& t, uint32_t& v)
{
    uint64_t w = 0;
    if (!parse_u64(t, w) || w > UINT32_MAX) return false;
    v = uint32_t(w);
    return true;
}
bool parse_dense(const std::string& t, size_t n, std::vector<uint64_t>& v)
{
    v.assign(n, 0);
    if (t.empty()) return true;
    size_t i = 0, pos = 0;
    for (;;) {
        const size_t e = t.find(',', pos);
        const std::string x = t.substr(pos, e == std::string::npos ? std::string::npos : e - pos);
        if (i >= n || (!x.empty() && !parse_u64(x, v[i]))) return false;
        ++i;
        if (e == std::string::npos) return true;
        pos = e + 1;
    }
}

std::string read_all(const std::string& path, bool& exists)
{
    std::FILE* f = std::fopen(path.c_str(), "rb");
    exists = f != nullptr;
    std::string all;
    if (!f) return all;
    char chunk[1 << 14];
    size_t r;
    while ((r = std::fread(chunk, 1, sizeof chunk, f)) > 0) all.append(chunk, r);
    std::fclose(f);
    return all;
}

// write + fsync a temp file, rename over path, fsync the directory
void write_durably(const std::string& path, const std::string& out)
{
    const std::string tmp = path + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw std::runtime_error("stats: cannot write " + tmp);
    bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    ok = std::fflush(f) == 0 && ok;
    ok = ::fsync(::fileno(f)) == 0 && ok;
    ok = std::fclose(f) == 0 && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        throw std::runtime_error("stats: cannot write " + path + " (state unchanged)");
    }
    const size_t sl = path.find_last_of('/');
    const std::string dir = sl == std::string::npos ? "." : (sl == 0 ? "/" : path.substr(0, sl));
    const int dfd = ::open(dir.c_str(), O_RDONLY | O_DIRECTORY | O_CLOEXEC);
    if (dfd >= 0) { ::fsync(dfd); ::close(dfd); }
}

struct FileLock {
    int fd = -1;
    explicit FileLock(const std::string& path)
    {
        fd = ::open(path.c_str(), O_RDWR | O_CREAT | O_CLOEXEC, 0644);
        if (fd < 0 || ::flock(fd, LOCK_EX) != 0) throw std::runtime_error("stats: cannot lock " + path);
    }
    ~FileLock() { if (fd >= 0) ::close(fd); }
    FileLock(const FileLock&) = delete;
    FileLock& operator=(const FileLock&) = delete;
};

}  // namespace

// ---------------------------------------------------------------- GapTotals

void GapTotals::walk_add(uint64_t g)
{
    const u128 g1 = g, g2 = g1 * g1;
    ++wn;
    wd[0] += g1; wd[1] += g2; wd[2] += g2 * g1; wd[3] += g2 * g2;
    samp.at(g / 2) += 1;
}

void GapTotals::walk_remove(uint64_t g)
{
    if (g >= GAP_TAIL || !wn || !samp.at(g / 2)) throw std::runtime_error("gaps: bad walk removal");
    const u128 g1 = g, g2 = g1 * g1;
    --wn;
    wd[0] -= g1; wd[1] -= g2; wd[2] -= g2 * g1; wd[3] -= g2 * g2;
    samp.at(g / 2) -= 1;
}

void GapTotals::exact_gap(uint64_t g, uint64_t lower)
{
    if (g >= GAP_TAIL) note_tail(uint32_t(g), 1, lower);
    note_max(g, lower);
}

void GapTotals::note_max(uint64_t g, uint64_t lower)
{
    if (g > gmax || (g == gmax && g && lower < gmax_lo)) { gmax = g; gmax_lo = lower; }
}

void GapTotals::note_tail(uint32_t g, uint64_t count, uint64_t lower)
{
    auto& t = tail[g];
    if (!t.first || lower < t.second) t.second = lower;
    t.first += count;
}

void GapTotals::merge(const GapTotals& o)
{
    n += o.n;
    d1 += o.d1;
    for (auto& t : o.tail) note_tail(t.first, t.second.first, t.second.second);
    note_max(o.gmax, o.gmax_lo);
    all_walked = all_walked && o.all_walked;
    wn += o.wn;
    for (int k = 0; k < 4; ++k) wd[k] += o.wd[k];
    for (size_t i = 0; i < samp.size(); ++i) samp[i] += o.samp[i];
}

// ---------------------------------------------------------------- GoldbachStats

namespace {
// The [15][np] count vectors (~145 KB each) of one statistics object come
// from a small per-thread pool and go back to it: allocated afresh for every
// run, the allocator returned their pages to the system at each free and
// every run faulted them in again (~50 us per vector); reused, they are only
// cleared.
thread_local std::vector<std::vector<uint64_t>> t_counts_pool;

void take_counts(std::vector<uint64_t>& v, size_t n)
{
    if (!t_counts_pool.empty()) {
        v = std::move(t_counts_pool.back());
        t_counts_pool.pop_back();
    }
    v.assign(n, 0);
}

void give_counts(std::vector<uint64_t>& v)
{
    if (v.capacity() != 0 && t_counts_pool.size() < 16) t_counts_pool.push_back(std::move(v));
}
}  // namespace

GoldbachStats::GoldbachStats(const std::vector<uint32_t>& cands, uint32_t tail_t, bool campaign)
    : cands_(cands), tail_t_(tail_t), campaign_(campaign)
{
    np_ = cands.size();
    pidx_.assign(cands.back() / 2 + 1, -1);
    for (size_t i = 0; i < np_; ++i) pidx_[cands[i] / 2] = int32_t(i);
    take_counts(samp_, 15 * np_);
    take_counts(tail_, 15 * np_);
    tail_agg_.assign(np_, 0);
    if (campaign_) {   // a run of its own fills these at its commit
        take_counts(cum_samp_, 15 * np_);
        take_counts(cum_tail_, 15 * np_);
        cum_tail_agg_.assign(np_, 0);
    }
    first_.assign(np_, 0);
    run_first_.assign(np_, 0);
    want_.assign(np_, 0);
    unknown_.assign(std::begin(kUnknownS), std::end(kUnknownS));
    state_path_ = env::path("GB_STATS_STATE", "goldbach_stats.state");
    log_path_ = env::path("GB_STATS_LOG", "goldbach_stats.log");
    rec_path_ = env::path("GB_RECORDS_FILE", "goldbach_records.txt");
    live_path_ = env::path("GB_STATS_LIVE", "goldbach_live.json");
    if (campaign_) {
        load_state();
        (void)load_gaps();
    }
    set_run_start(~0ull);
}

GoldbachStats::~GoldbachStats()
{
    give_counts(samp_);
    give_counts(tail_);
    give_counts(cum_samp_);
    give_counts(cum_tail_);
}

void GoldbachStats::set_run_start(uint64_t start)
{
    for (size_t i = 0; i < np_; ++i)
        want_[i] = cands_[i] >= tail_t_ && (!first_[i] || first_[i] > start);
}

uint64_t GoldbachStats::add_hist(bool sampled, const std::vector<uint64_t>& h)
{
    uint64_t n = 0;
    for (uint32_t cr = 0; cr < 15; ++cr)
        for (size_t i = 0; i < np_; ++i) {
            const uint64_t c = h[cr * np_ + i];
            if (!c) continue;
            const bool tail = cands_[i] >= tail_t_;
            if (sampled) {
                samp_[cr * np_ + i] += c;
                if (tail) tail_[cr * np_ + i] += c;
                n += c;
            } else if (tail) {
                tail_agg_[i] += c;
            }
        }
    evens_sampled_ += n;
    return n;
}

uint64_t GoldbachStats::add_hist_slots(bool sampled, const uint32_t* counts, size_t stride,
                                       const std::vector<uint16_t>* slot_pi)
{
    uint64_t n = 0;
    for (uint32_t cr = 0; cr < 15; ++cr) {
        const uint32_t* row = counts + cr * stride;
        const std::vector<uint16_t>& sp = slot_pi[cr];
        for (size_t k = 0; k < sp.size(); ++k) {
            const uint32_t c = row[k];
            if (!c) continue;
            const uint32_t i = sp[k];
            const bool tail = cands_[i] >= tail_t_;
            if (sampled) {
                samp_[cr * np_ + i] += c;
                if (tail) tail_[cr * np_ + i] += c;
                n += c;
            } else if (tail) {
                tail_agg_[i] += c;
            }
        }
    }
    evens_sampled_ += n;
    return n;
}

void GoldbachStats::add_first(uint64_t seg_start, const uint32_t* first)
{
    for (size_t i = 0; i < np_; ++i)
        if (first[i] != ~0u && want_[i]) note_first(uint32_t(i), seg_start + 2ull * first[i]);
}

void GoldbachStats::add_counts(const uint32_t cnt[8], const uint32_t tw[3])
{
    for (int i = 0; i < 8; ++i) primes_res_[i] += cnt[i];
    for (int i = 0; i < 3; ++i) twins_[i] += tw[i];
}

void GoldbachStats::begin_gaps(uint64_t start)
{
    gap_last_ = 0;             // prev_prime(start - 3), found when the first segment arrives
    gap_start_ = start;
}

void GoldbachStats::add_gap_segment(const SegmentGaps& s)
{
    GapTotals& r = gaps_;
    gap_last_walked_ = s.walked;
    r.all_walked = r.all_walked && s.walked;
    if (s.walked) walk_evens_ += s.evens;
    for (uint32_t i = 0; i < s.ntail; ++i) {
        const uint64_t up = s.q_first + s.tail[2 * i], g = s.tail[2 * i + 1];
        if (up >= gap_start_ + 1) r.exact_gap(g, up - g);
    }
    if (!s.primes) {
        if (s.pfirst != ~0u || s.gmax) throw std::runtime_error("gaps: a prime position in a segment without primes");
        return;
    }
    const uint64_t fq = s.q_first + s.pfirst, lq = s.q_first + s.plast;
    if (!gap_last_) gap_last_ = prev_prime(gap_start_ - 3);
    if (s.walked) {
        std::map<uint32_t, uint64_t> by_g;
        for (uint32_t i = 0; i < s.ntail; ++i) by_g[s.tail[2 * i + 1]] += 1;
        u128 m[4] = {};
        uint64_t hs = 0, ht = 0;
        bool tail_ok = true;
        for (uint32_t i = 1; i + 1 < GAP_BINS; ++i) {
            const u128 c = s.hist[i], g = 2 * i, g2 = g * g;
            hs += s.hist[i];
            m[0] += c * g; m[1] += c * g2; m[2] += c * g2 * g; m[3] += c * g2 * g2;
            if (2 * i >= GAP_TAIL) {
                ht += s.hist[i];
                const auto it = by_g.find(2 * i);
                tail_ok = tail_ok && s.hist[i] == (it == by_g.end() ? 0 : it->second);
            }
        }
        const uint64_t over = uint64_t(s.hist[GAP_BINS - 1]) + s.hist[GAP_BINS];
        const bool ok = s.hist[0] == 0 && over == 0 && hs + 1 == s.primes && m[0] == lq - fq && tail_ok &&
                        ht == s.ntail;
        if (!ok)
            throw std::runtime_error("gaps: cross-check failed (walk gaps " + std::to_string(hs) + ", primes " +
                                     std::to_string(s.primes) + ", sum " + u128_str(m[0]) + ", span " +
                                     std::to_string(lq - fq) + ", overflow " + std::to_string(over) + ", tails " +
                                     std::to_string(ht) + "/" + std::to_string(s.ntail) + (tail_ok ? "" : " by g") + ")");
        r.wn += hs;
        for (int k = 0; k < 4; ++k) r.wd[k] += m[k];
        for (uint32_t i = 1; i < GAP_BINS; ++i) r.samp[i] += s.hist[i];
        if (hs) {
            const uint64_t g = s.gmax >> 32, up = s.q_first + uint32_t(~uint32_t(s.gmax));
            if (!g || g >= 2 * GAP_BINS) throw std::runtime_error("gaps: walk maximum out of range");
            r.note_max(g, up - g);
        }
    }
    if (fq >= gap_start_ + 1) {
        r.exact_gap(fq - gap_last_, gap_last_);
        if (s.walked) r.walk_add(fq - gap_last_);
    }
    gap_last_ = lq;
}

void GoldbachStats::add_cpu(uint64_t n, uint64_t p, bool sampled)
{
    const uint32_t cr = uint32_t(n % 30) / 2;
    if (p <= cands_.back() && (p & 1) && pidx_[p / 2] >= 0) {
        const uint32_t pi = uint32_t(pidx_[p / 2]);
        if (sampled) { samp_[cr * np_ + pi] += 1; ++evens_sampled_; }
        if (p >= tail_t_) {
            if (sampled) tail_[cr * np_ + pi] += 1;
            else tail_agg_[pi] += 1;
            note_first(pi, n);
        }
    } else {
        beyond_.push_back({n, p});
        if (sampled) ++evens_sampled_;
        const bool rec = p > RECORD_P;
        record(n, p, rec ? "RECORD" : "BEYOND",
               rec ? "NEW RECORD: p(n) > 9781, the largest p(n) known below 4e18" : "p(n) beyond the candidate list");
    }
}

void GoldbachStats::finish_run(uint64_t start, uint64_t limit, uint64_t evens)
{
    start_ = start;
    limit_ = limit;
    evens_ = evens;
    // counted over [start - 3, limit - 3]; reported over [start + 1, limit - 1]
    for (uint64_t q : {start - 3, start - 1})
        if (is_prime_u64(q)) primes_res_[res_idx(q)] -= 1;
    if (is_prime_u64(limit - 1)) primes_res_[res_idx(limit - 1)] += 1;
    // twins counted for q in [start - 5, limit - 5]; reported q in [start + 1, limit - 3]
    for (uint64_t q : {start - 5, start - 3, start - 1})
        if (is_prime_u64(q) && is_prime_u64(q + 2)) twins_[twin_idx(q)] -= 1;
    if (is_prime_u64(limit - 3) && is_prime_u64(limit - 1)) twins_[twin_idx(limit - 3)] += 1;
    if (gap_start_ == start) {
        if (!gap_last_) gap_last_ = prev_prime(gap_start_ - 3);   // no segment had a prime
        gaps_.n = run_primes();
        if (limit >= start + 2) {
            if (is_prime_u64(start - 3) && is_prime_u64(start - 1)) gaps_.walk_remove(2);
            if (is_prime_u64(limit - 1)) {
                gaps_.exact_gap(limit - 1 - gap_last_, gap_last_);
                if (gap_last_walked_) gaps_.walk_add(limit - 1 - gap_last_);
            }
            uint64_t f = (start + 1) | 1;
            while (f <= limit - 1 && !is_prime_u64(f)) f += 2;
            if (f <= limit - 1) gaps_.d1 = prev_prime(limit) - prev_prime(f);
        }
    }
}

uint64_t GoldbachStats::run_primes() const { return std::accumulate(primes_res_, primes_res_ + 8, uint64_t(0)); }

uint64_t GoldbachStats::run_tail() const
{
    uint64_t t = beyond_.size();
    for (uint64_t v : tail_) t += v;
    for (uint64_t v : tail_agg_) t += v;
    return t;
}

uint64_t GoldbachStats::run_pmax() const
{
    uint64_t m = 0;
    for (auto& b : beyond_) m = std::max(m, b.second);
    if (m) return m;
    for (size_t i = np_; i-- > 0;) {
        if (tail_agg_[i]) return cands_[i];
        for (uint32_t cr = 0; cr < 15; ++cr)
            if (tail_[cr * np_ + i]) return cands_[i];
    }
    return 0;
}

void GoldbachStats::note_first(uint32_t pi, uint64_t n)
{
    const bool fresh = !run_first_[pi];
    if (fresh || n < run_first_[pi]) run_first_[pi] = n;
    const uint32_t p = cands_[pi];
    if (fresh && !first_[pi] && std::binary_search(unknown_.begin(), unknown_.end(), uint16_t(p)))
        record(n, p, "SP",
               "S(p) candidate: first even in this campaign with p(n) = p; S(p) was unknown (> 4e18) -- this IS "
               "S(p) if the campaign covers [4e18, n] without a gap");
}

void GoldbachStats::record(uint64_t n, uint64_t p, const char* kind, const char* what)
{
    // re-proof by the independent ascending scan before anything is written
    uint64_t ex = 0;
    for (uint64_t c = 3; c <= n / 2; c += 2)
        if (is_prime_u64(c) && is_prime_u64(n - c)) { ex = c; break; }
    char ts[32];
    const std::time_t t = std::time(nullptr);
    std::strftime(ts, sizeof ts, "%Y-%m-%dT%H:%M:%S", std::localtime(&t));
    char line[512];
    std::snprintf(line, sizeof line, "%s p(n)=%llu n=%llu q=%llu cpu_check=%s  %s\n", ts, (unsigned long long)p,
                  (unsigned long long)n, (unsigned long long)(n - p), ex == p ? "OK" : "MISMATCH", what);
    std::fprintf(stderr, "\n*** %s", line);
    if (std::FILE* f = std::fopen(rec_path_.c_str(), "ab")) { std::fputs(line, f); std::fclose(f); }
    events_.push_back({n, p});
    const char* hk = std::getenv("GB_RECORD_HOOK");
    std::string hook = hk ? hk : "";
    if (!hk)
        if (std::FILE* sf = std::fopen("scripts/notify_record.sh", "rb")) {
            std::fclose(sf);
            hook = "bash scripts/notify_record.sh";
        }
    if (!hook.empty()) {
        char cmd[1024];
        std::snprintf(cmd, sizeof cmd, "%s %s %llu %llu %llu %s >/dev/null 2>&1 &", hook.c_str(), kind,
                      (unsigned long long)p, (unsigned long long)n, (unsigned long long)(n - p),
                      ex == p ? "OK" : "MISMATCH");
        if (std::system(cmd) != 0) std::fprintf(stderr, "record hook failed to start: %s\n", hook.c_str());
    }
}

std::vector<uint64_t> GoldbachStats::by_prime(const std::vector<uint64_t>& cls, const std::vector<uint64_t>* agg) const
{
    std::vector<uint64_t> v(np_, 0);
    for (uint32_t cr = 0; cr < 15; ++cr)
        for (size_t i = 0; i < np_; ++i) v[i] += cls[cr * np_ + i];
    if (agg)
        for (size_t i = 0; i < np_; ++i) v[i] += (*agg)[i];
    return v;
}

std::string GoldbachStats::gap_fields() const
{
    const GapTotals& g = gaps_;
    std::string o = " gaps=" + std::to_string(g.n) + " gd1=" + u128_str(g.d1) + " gmax=" +
                    (g.gmax_exact() && g.gmax ? std::to_string(g.gmax) + ":" + std::to_string(g.gmax_lo) : "-") +
                    " gtail=";
    std::string t;
    for (auto& e : g.tail)
        t += (t.empty() ? "" : ";") + std::to_string(e.first) + ":" + std::to_string(e.second.first) + ":" +
             std::to_string(e.second.second);
    o += t.empty() ? "-" : t;
    o += " gwalk=" + std::to_string(walk_evens_) + ":" + std::to_string(g.wn) + " gsmom=";
    for (int k = 0; k < 4; ++k) o += (k ? "," : "") + u128_str(g.wd[k]);
    const std::string h = dense_list(g.samp);
    return o + " gsamp=" + (h.empty() ? std::string("-") : h);
}

std::string GoldbachStats::log_line(const char* version) const
{
    std::string line = "GBSTATS v2 version=" + std::string(version);
    line += " start=" + std::to_string(start_) + " limit=" + std::to_string(limit_) + " evens=" +
            std::to_string(evens_) + " evens_sampled=" + std::to_string(evens_sampled_) + " tail_t=" +
            std::to_string(tail_t_) + " tail_evens=" + std::to_string(run_tail()) + " pmax=" +
            std::to_string(run_pmax()) + " primes=" + std::to_string(run_primes()) + " twins=" +
            std::to_string(run_twins());
    line += " primes30=";
    for (int i = 0; i < 8; ++i) line += (i ? "," : "") + std::to_string(primes_res_[i]);
    line += " twins30=";
    for (int i = 0; i < 3; ++i) line += (i ? "," : "") + std::to_string(twins_[i]);
    std::string nf;
    for (size_t i = 0; i < np_; ++i)
        if (run_first_[i] && (!first_[i] || run_first_[i] < first_[i]))
            nf += (nf.empty() ? "" : ";") + std::to_string(cands_[i]) + ":" + std::to_string(run_first_[i]);
    line += " new_first=" + (nf.empty() ? std::string("-") : nf);
    std::string bey;
    for (auto& e : beyond_) bey += (bey.empty() ? "" : ";") + std::to_string(e.second) + ":" + std::to_string(e.first);
    line += " beyond=" + (bey.empty() ? std::string("-") : bey);
    if (gap_start_ == start_) line += gap_fields();
    size_t t0 = 0;
    while (t0 < np_ && cands_[t0] < tail_t_) ++t0;
    const std::string tl = t0 < np_ ? dense_list(by_prime(tail_, &tail_agg_), t0) : "";
    line += " tail=" + (tl.empty() ? std::string("-") : std::to_string(cands_[t0]) + ":" + tl) + "\n";
    return line;
}

void GoldbachStats::append_log(const std::string& line) const
{
    std::FILE* f = std::fopen(log_path_.c_str(), "ab");
    if (!f) { std::fprintf(stderr, "stats: cannot append to %s\n", log_path_.c_str()); return; }
    bool ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
    ok = std::fflush(f) == 0 && ok;
    ok = ::fsync(::fileno(f)) == 0 && ok;
    if (std::fclose(f) != 0 || !ok) std::fprintf(stderr, "stats: incomplete append to %s\n", log_path_.c_str());
}

bool GoldbachStats::log_has_gaps_in(uint64_t a, uint64_t b) const
{
    bool ex = false;
    const std::string all = read_all(log_path_, ex);
    size_t pos = 0;
    while (pos < all.size()) {
        size_t e = all.find('\n', pos);
        if (e == std::string::npos) e = all.size();
        const std::string ln = all.substr(pos, e - pos);
        pos = e + 1;
        if (ln.compare(0, 2, "GB") != 0 || ln.find(" gd1=") == std::string::npos) continue;
        const size_t s = ln.find(" start="), l = ln.find(" limit=");
        if (s != std::string::npos && l != std::string::npos &&
            std::strtoull(ln.c_str() + s + 7, nullptr, 10) <= b && a <= std::strtoull(ln.c_str() + l + 7, nullptr, 10))
            return true;
    }
    return false;
}

GoldbachStats::GapCampaign GoldbachStats::load_gaps() const
{
    GapCampaign gc;
    const std::string gp = state_path_ + ".gaps";
    bool ex = false;
    const std::string all = read_all(gp, ex);
    if (!ex) return gc;
    unsigned seen = 0;
    size_t pos = 0;
    const auto damaged = [&](const std::string& line) {
        return std::runtime_error("stats: gap state " + gp + " has a malformed record '" + line.substr(0, 80) +
                                  "'; refusing to merge into it");
    };
    while (pos < all.size()) {
        const size_t e = all.find('\n', pos);
        if (e == std::string::npos) break;
        const std::string buf = all.substr(pos, e - pos);
        pos = e + 1;
        if (buf.empty() || buf[0] == '#') continue;
        const std::vector<std::string> fl = split_spaces(buf);
        const std::string& k = fl[0];
        const size_t nf = fl.size();
        const auto once = [&](unsigned bit) {
            if (seen & bit) throw damaged(buf);
            seen |= bit;
        };
        bool ok = true;
        uint64_t a = 0, c = 0;
        if (k == "gaps_range" && nf == 3) {
            ok = parse_u64(fl[1], a) && parse_u64(fl[2], c) && a <= c;
            gc.covered.push_back({a, c});
            seen |= 1;
        } else if (k == "gaps_evens" && nf == 3) {
            once(2);
            ok = parse_u64(fl[1], gc.evens) && parse_u64(fl[2], gc.walked);
        } else if (k == "gaps" && nf == 3) {
            once(4);
            ok = parse_u64(fl[1], gc.g.n) && parse_u128(fl[2], gc.g.d1);
        } else if (k == "gaps_samp" && nf <= 2) {
            once(8);
            std::vector<uint64_t> v;
            ok = parse_dense(nf == 2 ? fl[1] : std::string(), GAP_BINS, v);
            std::copy(v.begin(), v.end(), gc.g.samp.begin());
        } else if (k == "gaps_max" && nf == 4) {
            once(16);
            ok = parse_u64(fl[1], a) && parse_u64(fl[2], c) && (fl[3] == "0" || fl[3] == "1");
            gc.g.gmax = a;
            gc.g.gmax_lo = c;
            gc.g.all_walked = fl[3] == "1" && a < GAP_TAIL;
        } else if (k == "gaps_walk" && nf == 6) {
            once(64);
            ok = parse_u64(fl[1], gc.g.wn);
            for (int i = 0; i < 4 && ok; ++i) ok = parse_u128(fl[2 + size_t(i)], gc.g.wd[i]);
        } else if (k == "gaps_tail" && nf == 4) {
            uint32_t g = 0;
            ok = parse_u32(fl[1], g) && parse_u64(fl[2], a) && parse_u64(fl[3], c);
            if (ok) gc.g.note_tail(g, a, c);
        } else if (k == "end" && nf == 1) {
            once(32);
        } else {
            ok = false;
        }
        if (!ok) throw damaged(buf);
    }
    if (seen != 127)
        throw std::runtime_error("stats: gap state " + gp + " is cut off or missing records; refusing to merge into it");
    merge_ranges(gc.covered);
    return gc;
}

bool GoldbachStats::gaps_cover(uint64_t a, uint64_t b) const
{
    for (auto& c : load_gaps().covered)
        if (c.first <= a && b <= c.second) return true;
    return false;
}

bool GoldbachStats::commit_gaps()
{
    GapCampaign gc = load_gaps();
    for (auto& c : gc.covered)
        if (start_ <= c.second && c.first <= limit_) {
            std::fprintf(stderr, "stats: gaps of [%llu, %llu] already in %s -- not recorded again\n",
                         (unsigned long long)start_, (unsigned long long)limit_, (state_path_ + ".gaps").c_str());
            return false;
        }
    gc.g.merge(gaps_);
    gc.evens += evens_;
    gc.walked += walk_evens_;
    gc.covered.push_back({start_, limit_});
    merge_ranges(gc.covered);
    std::string out =
        "# goldbach-l0 prime gap state v1 (exact; include/gstats.hpp GapStats)\n"
        "# gaps_range: lo hi with prime gaps recorded (upper prime in [lo + 1, hi - 1]);\n"
        "# gaps_evens <evens> <walked evens>; gaps <count> <sum g> (exact);\n"
        "# gaps_walk <count> <sum g> <sum g^2> <sum g^3> <sum g^4> and gaps_samp (histogram by g / 2): walked "
        "segments;\n"
        "# gaps_max <g> <lower> <1: exact, 0: largest seen>; gaps_tail <g> <count> <first lower> (g >= 700, exact)\n";
    for (auto& c : gc.covered) out += "gaps_range " + std::to_string(c.first) + " " + std::to_string(c.second) + "\n";
    out += "gaps_evens " + std::to_string(gc.evens) + " " + std::to_string(gc.walked) + "\n";
    out += "gaps " + std::to_string(gc.g.n) + " " + u128_str(gc.g.d1) + "\ngaps_walk " + std::to_string(gc.g.wn);
    for (int k = 0; k < 4; ++k) out += " " + u128_str(gc.g.wd[k]);
    out += "\ngaps_samp " + dense_list(gc.g.samp) + "\n";
    out += "gaps_max " + std::to_string(gc.g.gmax) + " " + std::to_string(gc.g.gmax_lo) + " " +
           (gc.g.gmax_exact() ? "1" : "0") + "\n";
    for (auto& t : gc.g.tail)
        out += "gaps_tail " + std::to_string(t.first) + " " + std::to_string(t.second.first) + " " +
               std::to_string(t.second.second) + "\n";
    out += "end\n";
    write_durably(state_path_ + ".gaps", out);
    return true;
}

void GoldbachStats::reset_campaign()
{
    std::fill(cum_samp_.begin(), cum_samp_.end(), 0);
    std::fill(cum_tail_.begin(), cum_tail_.end(), 0);
    std::fill(cum_tail_agg_.begin(), cum_tail_agg_.end(), 0);
    std::fill(first_.begin(), first_.end(), 0);
    covered_.clear();
    tot_beyond_.clear();
    tot_evens_ = tot_sampled_ = 0;
    for (auto& v : tot_primes_) v = 0;
    for (auto& v : tot_twins_) v = 0;
}

void GoldbachStats::load_state()
{
    bool ex = false;
    const std::string all = read_all(state_path_, ex);
    if (!ex) return;
    if (all.empty() || all.back() != '\n')
        throw std::runtime_error("stats: state " + state_path_ +
                                 " is empty or cut off (truncated write?); refusing to merge into it");
    uint32_t seen_samp = 0, seen_tail = 0;
    bool seen_t = false, seen_ev = false, seen_p = false, seen_tw = false, seen_agg = false;
    const auto damaged = [&](const std::string& line) {
        return std::runtime_error("stats: state " + state_path_ + " has a malformed record '" + line.substr(0, 80) +
                                  "' (damaged); refusing to merge into it");
    };
    const auto once = [&](bool& seen, const std::string& line) {
        if (seen) throw damaged(line);
        seen = true;
    };
    size_t pos = 0;
    while (pos < all.size()) {
        size_t e = all.find('\n', pos);
        if (e == std::string::npos) e = all.size();
        const std::string buf = all.substr(pos, e - pos);
        pos = e + 1;
        if (buf.empty() || buf[0] == '#') continue;
        const std::vector<std::string> fl = split_spaces(buf);
        const std::string& k = fl[0];
        const size_t nf = fl.size();
        bool ok = true;
        uint64_t a = 0, c = 0;
        uint32_t p = 0, cr = 0;
        if (k == "range" && nf == 3) {
            ok = parse_u64(fl[1], a) && parse_u64(fl[2], c) && a <= c;
            covered_.push_back({a, c});
        } else if (k == "tail_t" && nf == 2) {
            once(seen_t, buf);
            ok = parse_u32(fl[1], p);
            if (ok && p != tail_t_)
                throw std::runtime_error("stats: state " + state_path_ + " was recorded with tail_t " +
                                         std::to_string(p) + ", this run uses " + std::to_string(tail_t_));
        } else if (k == "evens" && nf == 3) {
            once(seen_ev, buf);
            ok = parse_u64(fl[1], tot_evens_) && parse_u64(fl[2], tot_sampled_);
        } else if (k == "first" && nf == 3) {
            ok = parse_u32(fl[1], p) && parse_u64(fl[2], a) && p <= cands_.back() && (p & 1u) && pidx_[p / 2] >= 0;
            if (ok) first_[size_t(pidx_[p / 2])] = a;
        } else if (k == "beyond" && nf == 3) {
            ok = parse_u64(fl[1], c) && parse_u64(fl[2], a);
            tot_beyond_.push_back({a, c});
        } else if (k == "primes30" && nf == 9) {
            once(seen_p, buf);
            for (size_t i = 0; i < 8 && ok; ++i) ok = parse_u64(fl[1 + i], tot_primes_[i]);
        } else if (k == "twins30" && nf == 4) {
            once(seen_tw, buf);
            for (size_t i = 0; i < 3 && ok; ++i) ok = parse_u64(fl[1 + i], tot_twins_[i]);
        } else if ((k == "sampled" || k == "tailcls") && nf == 3) {
            uint32_t& seen = k == "sampled" ? seen_samp : seen_tail;
            std::vector<uint64_t> v;
            ok = parse_u32(fl[1], cr) && cr < 15 && !((seen >> cr) & 1u) && parse_dense(fl[2], np_, v);
            if (ok) {
                std::copy(v.begin(), v.end(), (k == "sampled" ? cum_samp_ : cum_tail_).begin() + cr * np_);
                seen |= 1u << cr;
            }
        } else if (k == "tailagg" && nf == 2) {
            once(seen_agg, buf);
            std::vector<uint64_t> v;
            ok = parse_dense(fl[1], np_, v);
            if (ok) cum_tail_agg_ = v;
        } else {
            ok = false;
        }
        if (!ok) throw damaged(buf);
    }
    if (!(seen_t && seen_ev && seen_p && seen_tw && seen_agg && seen_samp == 0x7FFFu && seen_tail == 0x7FFFu))
        throw std::runtime_error("stats: state " + state_path_ +
                                 " is missing records (truncated or damaged); refusing to merge into it");
    merge_ranges(covered_);
}

void GoldbachStats::save_state() const { write_durably(state_path_, state_text()); }

std::string GoldbachStats::state_text() const
{
    std::string out =
        "# goldbach-l0 statistics state v2 (exact; include/gstats.hpp)\n"
        "# lists: counts by odd-prime index (3, 5, 7, 11, ...), empty = 0\n"
        "# sampled <n mod 30 / 2>: p(n) histogram of the sampled segments\n"
        "# tailcls <n mod 30 / 2> / tailagg: p(n) >= tail_t, all evens (tailagg: class not kept)\n"
        "# first <p> <n>: first even of the campaign with p(n) = p (tail p only)\n";
    out += "tail_t " + std::to_string(tail_t_) + "\n";
    for (auto& c : covered_) out += "range " + std::to_string(c.first) + " " + std::to_string(c.second) + "\n";
    out += "evens " + std::to_string(tot_evens_) + " " + std::to_string(tot_sampled_) + "\n";
    out += "primes30";
    for (int i = 0; i < 8; ++i) out += " " + std::to_string(tot_primes_[i]);
    out += "\ntwins30";
    for (int i = 0; i < 3; ++i) out += " " + std::to_string(tot_twins_[i]);
    out += "\n";
    for (size_t i = 0; i < np_; ++i)
        if (first_[i]) out += "first " + std::to_string(cands_[i]) + " " + std::to_string(first_[i]) + "\n";
    for (auto& e : tot_beyond_) out += "beyond " + std::to_string(e.second) + " " + std::to_string(e.first) + "\n";
    for (uint32_t cr = 0; cr < 15; ++cr) {
        std::vector<uint64_t> v(cum_samp_.begin() + cr * np_, cum_samp_.begin() + (cr + 1) * np_);
        out += "sampled " + std::to_string(cr) + " " + dense_list(v) + "\n";
    }
    for (uint32_t cr = 0; cr < 15; ++cr) {
        std::vector<uint64_t> v(cum_tail_.begin() + cr * np_, cum_tail_.begin() + (cr + 1) * np_);
        out += "tailcls " + std::to_string(cr) + " " + dense_list(v) + "\n";
    }
    out += "tailagg " + dense_list(cum_tail_agg_) + "\n";
    return out;
}

std::string GoldbachStats::memory_commit(uint64_t sampled_total)
{
    if (campaign_) throw std::runtime_error("stats: memory_commit on a campaign run");
    if (evens_sampled_ != sampled_total)
        throw std::runtime_error("stats: " + std::to_string(evens_sampled_) + " sampled evens recorded, " +
                                 std::to_string(sampled_total) + " in the sampled segments");
    cum_samp_ = samp_;     // a fresh campaign: its totals are this run's
    cum_tail_ = tail_;
    cum_tail_agg_ = tail_agg_;
    for (size_t i = 0; i < np_; ++i)
        if (run_first_[i] && (!first_[i] || run_first_[i] < first_[i])) first_[i] = run_first_[i];
    for (int i = 0; i < 8; ++i) tot_primes_[i] += primes_res_[i];
    for (int i = 0; i < 3; ++i) tot_twins_[i] += twins_[i];
    tot_evens_ += evens_;
    tot_sampled_ += evens_sampled_;
    for (auto& b : beyond_) tot_beyond_.push_back(b);
    covered_.push_back({start_, limit_});
    return state_text();
}

void GoldbachStats::commit(const char* version, uint64_t sampled_total, const std::string* run_line)
{
    if (!campaign_) throw std::runtime_error("stats: commit on a run without campaign files");
    if (evens_sampled_ != sampled_total)
        throw std::runtime_error("stats: " + std::to_string(evens_sampled_) + " sampled evens recorded, " +
                                 std::to_string(sampled_total) + " in the sampled segments (statistics not written)");
    const FileLock lock(state_path_ + ".lock");
    // the log line names first occurrences new to the campaign: the caller's
    // copy holds only if the state read here has the first occurrences the
    // run started from
    const std::vector<uint64_t> first_before = run_line ? first_ : std::vector<uint64_t>();
    reset_campaign();
    load_state();
    if (run_line && first_ != first_before) run_line = nullptr;
    for (auto& c : covered_)
        if (start_ <= c.second && c.first <= limit_) {
            std::fprintf(stderr, "stats: [%llu, %llu] overlaps the covered range [%llu, %llu] of %s -- not recorded again\n",
                         (unsigned long long)start_, (unsigned long long)limit_, (unsigned long long)c.first,
                         (unsigned long long)c.second, state_path_.c_str());
            if (gap_start_ == start_) {
                append_log("GBPSAMP v1 version=" + std::string(version) + " start=" + std::to_string(start_) +
                           " limit=" + std::to_string(limit_) + " evens=" + std::to_string(evens_) +
                           " psev=" + std::to_string(evens_sampled_) + " psamp=" + dense_list(by_prime(samp_, nullptr)) +
                           "\n");
                const std::string row = "GBGAPS v1 version=" + std::string(version) + " start=" + std::to_string(start_) +
                                        " limit=" + std::to_string(limit_) + " evens=" + std::to_string(evens_) +
                                        gap_fields() + "\n";
                if (commit_gaps())
                    append_log(row);
                else if (gaps_cover(start_, limit_) && !log_has_gaps_in(start_, limit_)) {
                    std::fprintf(stderr, "stats: %s has no gap row for [%llu, %llu] -- written now\n",
                                 log_path_.c_str(), (unsigned long long)start_, (unsigned long long)limit_);
                    append_log(row);
                }
            }
            return;
        }
    const std::string line = run_line ? *run_line : log_line(version);
    for (size_t i = 0; i < samp_.size(); ++i) { cum_samp_[i] += samp_[i]; cum_tail_[i] += tail_[i]; }
    for (size_t i = 0; i < np_; ++i) cum_tail_agg_[i] += tail_agg_[i];
    for (size_t i = 0; i < np_; ++i)
        if (run_first_[i] && (!first_[i] || run_first_[i] < first_[i])) first_[i] = run_first_[i];
    for (int i = 0; i < 8; ++i) tot_primes_[i] += primes_res_[i];
    for (int i = 0; i < 3; ++i) tot_twins_[i] += twins_[i];
    tot_evens_ += evens_;
    tot_sampled_ += evens_sampled_;
    if (gap_start_ == start_) commit_gaps();
    for (auto& b : beyond_) tot_beyond_.push_back(b);
    covered_.push_back({start_, limit_});
    merge_ranges(covered_);
    save_state();
    append_log(line);
}

void GoldbachStats::write_live(uint64_t start, uint64_t limit, uint64_t done, uint64_t total, double elapsed, bool force)
{
    if (!campaign_) return;
    const long long now = (long long)std::time(nullptr);
    if (!force && now - live_t_ < 2) return;
    live_t_ = now;
    uint64_t pm = 0, pmn = 0, ge[3] = {};
    for (size_t i = 0; i < np_; ++i)
        if (run_first_[i] && cands_[i] > pm) { pm = cands_[i]; pmn = run_first_[i]; }
    for (auto& b : beyond_)
        if (b.second > pm) { pm = b.second; pmn = b.first; }
    const uint64_t lims[3] = {8000, 9000, RECORD_P + 1};
    for (size_t i = 0; i < np_; ++i) {
        uint64_t c = tail_agg_[i];
        for (uint32_t cr = 0; cr < 15; ++cr) c += tail_[cr * np_ + i];
        for (int k = 0; k < 3; ++k)
            if (cands_[i] >= lims[k]) ge[k] += c;
    }
    ge[2] += beyond_.size();
    char b[512];
    std::snprintf(b, sizeof b,
                  "{\"time\": %lld, \"start\": %llu, \"limit\": %llu, \"evens_done\": %llu, \"evens_total\": %llu, "
                  "\"elapsed_s\": %.1f, \"tail_t\": %u, \"tail_evens\": %llu, \"pmax\": %llu, \"pmax_n\": %llu, "
                  "\"ge8000\": %llu, \"ge9000\": %llu, \"gt9781\": %llu, \"primes\": %llu, \"twins\": %llu, "
                  "\"events\": [",
                  now, (unsigned long long)start, (unsigned long long)limit, (unsigned long long)done,
                  (unsigned long long)total, elapsed, tail_t_, (unsigned long long)run_tail(), (unsigned long long)pm,
                  (unsigned long long)pmn, (unsigned long long)ge[0], (unsigned long long)ge[1],
                  (unsigned long long)ge[2], (unsigned long long)run_primes(), (unsigned long long)run_twins());
    std::string o = b;
    for (size_t i = 0; i < events_.size(); ++i) {
        std::snprintf(b, sizeof b, "%s[%llu, \"%llu\"]", i ? ", " : "", (unsigned long long)events_[i].second,
                      (unsigned long long)events_[i].first);
        o += b;
    }
    o += "]}\n";
    const std::string tmp = live_path_ + ".tmp";
    if (std::FILE* f = std::fopen(tmp.c_str(), "wb")) {
        const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
        if (std::fclose(f) == 0 && ok) std::rename(tmp.c_str(), live_path_.c_str());
    }
}

}  // namespace gb
