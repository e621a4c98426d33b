// Exact statistics and campaign files (contracts, field meanings and the
// design in gstats.hpp).
#include "gstats.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <string_view>

#include "env.hpp"
#include "primes.hpp"
#include "textio.hpp"

namespace gb {

std::string u128_str(u128 v) { return textio::dec(v); }

namespace {

using Ranges = std::vector<std::pair<uint64_t, uint64_t>>;
using textio::count_list;
using textio::dec;

// Primes p <= 9781 for which no even n <= 4e18 has p(n) = p (the unknown
// entries of T. Oliveira e Silva's published S(p) table, as the reference
// lists them at include/gstats.hpp:127-136): the first even of a campaign
// with such a p(n) is announced as an S(p) candidate.
constexpr uint16_t kSpUnknown[] = {
    8423, 8429, 8537, 8627, 8669, 8713, 8741, 8753, 8783, 8803, 8807, 8837, 8839, 8849, 8861,
    8863, 8867, 8887, 8893, 8923, 8963, 8969, 8999, 9007, 9011, 9013, 9029, 9041, 9043, 9049,
    9059, 9067, 9103, 9109, 9127, 9133, 9137, 9151, 9157, 9173, 9181, 9187, 9199, 9209, 9221,
    9227, 9239, 9241, 9257, 9277, 9281, 9283, 9293, 9311, 9319, 9323, 9337, 9343, 9349, 9371,
    9377, 9391, 9397, 9403, 9413, 9419, 9421, 9431, 9433, 9437, 9439, 9461, 9463, 9467, 9473,
    9479, 9491, 9497, 9511, 9521, 9533, 9539, 9547, 9551, 9587, 9601, 9613, 9619, 9623, 9631,
    9643, 9649, 9661, 9677, 9679, 9689, 9697, 9719, 9721, 9733, 9739, 9743, 9749, 9767, 9769};

bool sp_unknown(uint32_t p) { return std::binary_search(std::begin(kSpUnknown), std::end(kSpUnknown), uint16_t(p)); }

// position of q mod 30 among the eight unit residues 1 7 11 13 17 19 23 29
// (the primes30 order); -1 for the other residues
int unit_slot(uint64_t q)
{
    constexpr uint32_t kUnits = (1u << 1) | (1u << 7) | (1u << 11) | (1u << 13) | (1u << 17) | (1u << 19) |
                                (1u << 23) | (1u << 29);
    const uint32_t r = uint32_t(q % 30);
    if (!((kUnits >> r) & 1u)) return -1;
    return __builtin_popcount(kUnits & ((1u << r) - 1));
}

// unit_slot of a prime the counts hold (every prime above 5)
size_t prime_slot(uint64_t q)
{
    const int s = unit_slot(q);
    if (s < 0) throw std::runtime_error("stats: prime " + std::to_string(q) + " outside the eight classes");
    return size_t(s);
}

// twin pair kind of its lower prime: 11 mod 30 -> 0, 17 -> 1, 29 -> 2
int twin_slot(uint64_t q)
{
    switch (q % 30) {
    case 11: return 0;
    case 17: return 1;
    default: return 2;
    }
}

// sorted, and ranges that overlap or touch (the next start at most 2 above
// the previous limit: consecutive evens) joined
void normalize(Ranges& r)
{
    std::sort(r.begin(), r.end());
    size_t w = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        if (w && r[i].first <= r[w - 1].second + 2) {
            r[w - 1].second = std::max(r[w - 1].second, r[i].second);
        } else {
            r[w++] = r[i];
        }
    }
    r.resize(w);
}

bool intersects(const Ranges& r, uint64_t a, uint64_t b, std::pair<uint64_t, uint64_t>* hit = nullptr)
{
    for (const auto& c : r)
        if (a <= c.second && c.first <= b) {
            if (hit) *hit = c;
            return true;
        }
    return false;
}

// ---- line-rule reader of the state and gap files -------------------------

using Words = std::vector<std::string_view>;

// One record kind of a line-oriented file: lines whose first word is `key`
// and that have between min_words and max_words words (the key included).
// `times`: 1 = exactly once, 0 = any number, 15 = once per even class (the
// class is the second word). `take` stores the record, false if malformed.
struct Rule {
    std::string_view key;
    size_t min_words, max_words;
    int times;
    std::function<bool(const Words&)> take;
};

// Applies `rules` to every line of `text` (blank lines and '#' comments
// skipped); a line no rule takes, a malformed record or a repeated one is a
// damaged file, a missing record an incomplete one: both throw with the
// file's name, nothing having been merged.
void read_by_rules(const std::string& text, const std::string& what, const std::vector<Rule>& rules)
{
    const auto damaged = [&](std::string_view line) {
        return std::runtime_error("stats: " + what + " has a malformed record '" +
                                  std::string(line.substr(0, 80)) + "'; refusing to merge into it");
    };
    std::vector<uint32_t> seen(rules.size(), 0);
    size_t at = 0;
    while (at < text.size()) {
        const size_t nl = text.find('\n', at);
        if (nl == std::string::npos)
            throw std::runtime_error("stats: " + what + " ends in a cut-off line; refusing to merge into it");
        const std::string_view line(text.data() + at, nl - at);
        at = nl + 1;
        if (line.empty() || line.front() == '#') continue;
        const Words w = textio::words(line);
        size_t k = 0;
        while (k < rules.size() && rules[k].key != w[0]) ++k;
        if (k == rules.size() || w.size() < rules[k].min_words || w.size() > rules[k].max_words) throw damaged(line);
        const Rule& r = rules[k];
        uint32_t bit = 1;
        if (r.times == 15) {
            uint32_t cr = 0;
            if (!textio::parse(w[1], cr) || cr >= 15) throw damaged(line);
            bit = 1u << cr;
        }
        if (r.times && (seen[k] & bit)) throw damaged(line);
        seen[k] |= bit;
        if (!r.take(w)) throw damaged(line);
    }
    for (size_t k = 0; k < rules.size(); ++k) {
        const uint32_t want = rules[k].times == 15 ? 0x7FFFu : rules[k].times == 1 ? 1u : 0u;
        if ((seen[k] & want) != want)
            throw std::runtime_error("stats: " + what + " lacks its '" + std::string(rules[k].key) +
                                     "' records (cut off or damaged); refusing to merge into it");
    }
}

// ---- count vectors --------------------------------------------------------
// The [15][np] count vectors (~145 KB each) of a statistics object come from
// a small per-thread free list: fresh for every run, the allocator gave their
// pages back to the system at each free and every run faulted them in again.
struct VectorShelf {
    std::vector<std::vector<uint64_t>> free;
    ~VectorShelf() = default;
};
thread_local VectorShelf t_shelf;

void take_shelved(std::vector<uint64_t>& v)
{
    if (!t_shelf.free.empty()) {
        v.swap(t_shelf.free.back());
        t_shelf.free.pop_back();
    }
}

void borrow(std::vector<uint64_t>& v, size_t n)
{
    take_shelved(v);
    v.assign(n, 0);
}

// v = src, in a shelved vector when one is free
void borrow_copy(std::vector<uint64_t>& v, const std::vector<uint64_t>& src)
{
    take_shelved(v);
    v.assign(src.begin(), src.end());
}

void give_back(std::vector<uint64_t>& v)
{
    if (v.capacity() && t_shelf.free.size() < 16) {
        t_shelf.free.emplace_back();
        t_shelf.free.back().swap(v);
    }
}

}  // namespace

// ---------------------------------------------------------------- GapBook

void GapBook::walk(uint64_t g, int sign)
{
    const size_t bin = g / 2;
    if (bin >= walk_hist.size()) throw std::runtime_error("gaps: walked gap " + std::to_string(g) + " beyond the bins");
    if (sign < 0 && (!walk_count || !walk_hist[bin] || g >= GAP_TAIL))
        throw std::runtime_error("gaps: taking out a gap the walk does not hold");
    const u128 p1 = g, p2 = p1 * p1, terms[4] = {p1, p2, p2 * p1, p2 * p2};
    for (int k = 0; k < 4; ++k) moment[k] = sign > 0 ? moment[k] + terms[k] : moment[k] - terms[k];
    walk_count += sign > 0 ? 1 : uint64_t(-1);
    walk_hist[bin] += sign > 0 ? 1 : uint64_t(-1);
}

void GapBook::offer_top(uint64_t g, uint64_t lower)
{
    const bool larger = g > top_g, earlier_tie = g == top_g && g != 0 && lower < top_lower;
    if (larger || earlier_tie) {
        top_g = g;
        top_lower = lower;
    }
}

void GapBook::add_tail(uint32_t g, uint64_t n, uint64_t lower)
{
    Tail& t = tail[g];
    t.first_lower = t.count ? std::min(t.first_lower, lower) : lower;
    t.count += n;
}

void GapBook::exact(uint64_t g, uint64_t lower)
{
    if (g >= GAP_TAIL) add_tail(uint32_t(g), 1, lower);
    offer_top(g, lower);
}

void GapBook::absorb(const GapBook& o)
{
    count += o.count;
    span += o.span;
    for (const auto& [g, t] : o.tail) add_tail(g, t.count, t.first_lower);
    offer_top(o.top_g, o.top_lower);
    all_walked = all_walked && o.all_walked;
    walk_count += o.walk_count;
    for (int k = 0; k < 4; ++k) moment[k] += o.moment[k];
    for (size_t i = 0; i < walk_hist.size(); ++i) walk_hist[i] += o.walk_hist[i];
}

// ---------------------------------------------------------------- Tally

void Tally::size_for(size_t np, bool by_class)
{
    if (by_class) {
        borrow(hist, 15 * np);
        borrow(tail_cls, 15 * np);
    }
    tail_flat.assign(np, 0);
    first.assign(np, 0);
    zero = true;
}

void Tally::size_by_class(size_t np)
{
    if (hist.size() == 15 * np) return;
    borrow(hist, 15 * np);
    borrow(tail_cls, 15 * np);
}

void Tally::clear()
{
    evens = sampled = 0;
    std::fill(std::begin(primes), std::end(primes), 0);
    std::fill(std::begin(twins), std::end(twins), 0);
    if (!zero)
        for (auto* v : {&hist, &tail_cls, &tail_flat, &first}) std::fill(v->begin(), v->end(), 0);
    zero = true;
    beyond.clear();
    ranges.clear();
}

void Tally::absorb(const Tally& run)
{
    zero = false;
    evens += run.evens;
    sampled += run.sampled;
    for (int i = 0; i < 8; ++i) primes[i] += run.primes[i];
    for (int i = 0; i < 3; ++i) twins[i] += run.twins[i];
    if (hist.empty()) {   // no class counts held yet (a fresh campaign): the run's
        borrow_copy(hist, run.hist);
        borrow_copy(tail_cls, run.tail_cls);
    } else {
        std::transform(hist.begin(), hist.end(), run.hist.begin(), hist.begin(), std::plus<>());
        std::transform(tail_cls.begin(), tail_cls.end(), run.tail_cls.begin(), tail_cls.begin(), std::plus<>());
    }
    std::transform(tail_flat.begin(), tail_flat.end(), run.tail_flat.begin(), tail_flat.begin(), std::plus<>());
    for (size_t i = 0; i < first.size(); ++i)
        if (run.first[i] && (!first[i] || run.first[i] < first[i])) first[i] = run.first[i];
    beyond.insert(beyond.end(), run.beyond.begin(), run.beyond.end());
    ranges.insert(ranges.end(), run.ranges.begin(), run.ranges.end());
    normalize(ranges);
}

// ---------------------------------------------------------------- GoldbachStats

GoldbachStats::GoldbachStats(const std::vector<uint32_t>& cands, uint32_t tail_t, bool campaign)
    : cands_(cands), tail_t_(tail_t), campaign_(campaign)
{
    np_ = cands.size();
    tail0_ = size_t(std::lower_bound(cands.begin(), cands.end(), tail_t) - cands.begin());
    index_of_.assign(cands.back() / 2 + 1, -1);
    for (size_t i = 0; i < np_; ++i) index_of_[cands[i] / 2] = int32_t(i);
    run_.size_for(np_, true);
    camp_.size_for(np_, false);   // its class counts only when a state file holds some (read_state) or a run is absorbed
    want_.assign(np_, 0);
    state_path_ = env::path("GB_STATS_STATE", "goldbach_stats.state");
    log_path_ = env::path("GB_STATS_LOG", "goldbach_stats.log");
    rec_path_ = env::path("GB_RECORDS_FILE", "goldbach_records.txt");
    live_path_ = env::path("GB_STATS_LIVE", "goldbach_live.json");
    if (campaign_) {   // damaged campaign files fail here, before any GPU work
        read_state(camp_);
        (void)read_gap_file();
    }
    set_run_start(~0ull);
}

GoldbachStats::~GoldbachStats()
{
    for (Tally* t : {&run_, &camp_}) {
        give_back(t->hist);
        give_back(t->tail_cls);
    }
}

void GoldbachStats::set_run_start(uint64_t start)
{
    for (size_t i = 0; i < np_; ++i) {
        const uint64_t f = camp_.first[i];
        want_[i] = i >= tail0_ && (f == 0 || f > start);
    }
}

void GoldbachStats::count_hit(bool sampled, uint32_t cr, uint32_t pi, uint64_t c)
{
    const size_t at = size_t(cr) * np_ + pi;
    const bool in_tail = pi >= tail0_;
    if (sampled) {
        run_.hist[at] += c;
        run_.sampled += c;
        if (in_tail) run_.tail_cls[at] += c;
    } else if (in_tail) {
        run_.tail_flat[pi] += c;
    }
}

uint64_t GoldbachStats::add_hist_slots(bool sampled, const uint32_t* counts, size_t stride,
                                       const std::vector<uint16_t>* slot_pi)
{
    // count_hit for every slot, in one pass per class
    run_.zero = false;
    const uint64_t before = run_.sampled;
    for (uint32_t cr = 0; cr < 15; ++cr) {
        const uint32_t* row = counts + size_t(cr) * stride;
        const std::vector<uint16_t>& pis = slot_pi[cr];
        const size_t n = pis.size();
        uint64_t* tc = run_.tail_cls.data() + size_t(cr) * np_;
        if (sampled) {
            uint64_t* h = run_.hist.data() + size_t(cr) * np_;
            uint64_t sum = 0;
            for (size_t k = 0; k < n; ++k) {
                const uint32_t c = row[k], pi = pis[k];
                h[pi] += c;
                sum += c;
                if (pi >= tail0_) tc[pi] += c;
            }
            run_.sampled += sum;
        } else {
            for (size_t k = 0; k < n; ++k)
                if (pis[k] >= tail0_) run_.tail_flat[pis[k]] += row[k];
        }
    }
    return run_.sampled - before;
}

void GoldbachStats::add_first(uint64_t seg_start, const uint32_t* first)
{
    run_.zero = false;
    for (size_t i = tail0_; i < np_; ++i)
        if (first[i] != ~0u && want_[i]) note_first(uint32_t(i), seg_start + 2 * uint64_t(first[i]));
}

void GoldbachStats::add_counts(const uint32_t cnt[8], const uint32_t tw[3])
{
    for (int i = 0; i < 8; ++i) run_.primes[i] += cnt[i];
    for (int i = 0; i < 3; ++i) run_.twins[i] += tw[i];
}

void GoldbachStats::begin_gaps(uint64_t start)
{
    gaps_from_ = start;
    last_prime_ = before_start_ = 0;   // prev_prime(start - 3), looked up when the first segment arrives
}

void GoldbachStats::add_gap_segment(const SegmentGaps& s)
{
    GapBook& b = gaps_;
    last_walked_ = s.walked;
    b.all_walked = b.all_walked && s.walked;
    if (s.walked) walked_evens_ += s.evens;
    // the device's inner tail gaps, those whose upper prime is in the run
    for (uint32_t i = 0; i < s.ntail; ++i) {
        const uint64_t upper = s.q_first + s.tail[2 * i], g = s.tail[2 * i + 1];
        if (upper > gaps_from_) b.exact(g, upper - g);
    }
    if (s.primes == 0) {
        if (s.pfirst != ~0u || s.gmax != 0)
            throw std::runtime_error("gaps: a segment without primes reports a prime position");
        return;
    }
    const uint64_t first_q = s.q_first + s.pfirst, last_q = s.q_first + s.plast;
    if (last_prime_ == 0) last_prime_ = before_start_ = prev_prime(gaps_from_ - 3);
    if (s.walked) {
        // the walked histogram must agree with the segment's own figures:
        // one gap fewer than primes, the gaps summing to the span between
        // its first and last prime, no gap beyond the bins, and every tail
        // bin matching the listed tail gaps
        std::map<uint32_t, uint64_t> listed;
        for (uint32_t i = 0; i < s.ntail; ++i) ++listed[s.tail[2 * i + 1]];
        uint64_t gaps = 0, tails = 0;
        u128 sum[4] = {};
        bool tails_agree = true;
        for (uint32_t bin = 1; bin + 1 < GAP_BINS; ++bin) {
            const uint64_t c = s.hist[bin];
            const u128 g = 2 * bin, g2 = g * g;
            gaps += c;
            sum[0] += c * g;
            sum[1] += c * g2;
            sum[2] += c * g2 * g;
            sum[3] += c * g2 * g2;
            if (2 * bin >= GAP_TAIL) {
                tails += c;
                const auto it = listed.find(2 * bin);
                tails_agree = tails_agree && c == (it == listed.end() ? 0 : it->second);
            }
        }
        const uint64_t beyond_bins = uint64_t(s.hist[GAP_BINS - 1]) + s.hist[GAP_BINS];
        if (s.hist[0] || beyond_bins || gaps + 1 != s.primes || sum[0] != last_q - first_q || !tails_agree ||
            tails != s.ntail)
            throw std::runtime_error("gaps: the walk does not add up (gaps " + std::to_string(gaps) + " for " +
                                     std::to_string(s.primes) + " primes, sum " + dec(sum[0]) + " over a span of " +
                                     std::to_string(last_q - first_q) + ", " + std::to_string(beyond_bins) +
                                     " beyond the bins, " + std::to_string(tails) + " tail gaps of " +
                                     std::to_string(s.ntail) + " listed" + (tails_agree ? "" : ", by size") + ")");
        b.walk_count += gaps;
        for (int k = 0; k < 4; ++k) b.moment[k] += sum[k];
        for (uint32_t bin = 1; bin < GAP_BINS; ++bin) b.walk_hist[bin] += s.hist[bin];
        if (gaps) {
            const uint64_t g = s.gmax >> 32, upper = s.q_first + uint32_t(~uint32_t(s.gmax));
            if (g == 0 || g >= 2 * GAP_BINS) throw std::runtime_error("gaps: the walk's largest gap is out of range");
            b.offer_top(g, upper - g);
        }
    }
    // the gap into the segment's first prime, from the last prime before it
    if (first_q > gaps_from_) {
        const uint64_t g = first_q - last_prime_;
        b.exact(g, last_prime_);
        if (s.walked) b.walk(g, +1);
    }
    last_prime_ = last_q;
}

void GoldbachStats::add_cpu(uint64_t n, uint64_t p, bool sampled)
{
    run_.zero = false;
    const int32_t pi = (p & 1) && p <= cands_.back() ? index_of_[p / 2] : -1;
    if (pi >= 0) {
        count_hit(sampled, uint32_t(n % 30) / 2, uint32_t(pi), 1);
        if (uint32_t(pi) >= tail0_) note_first(uint32_t(pi), n);
        return;
    }
    run_.beyond.push_back({n, p});
    if (sampled) ++run_.sampled;
    if (p > RECORD_P)
        announce(n, p, "RECORD", "NEW RECORD: p(n) > 9781, the largest p(n) known below 4e18");
    else
        announce(n, p, "BEYOND", "p(n) beyond the candidate list");
}

void GoldbachStats::finish_run(uint64_t start, uint64_t limit, uint64_t evens)
{
    start_ = start;
    limit_ = limit;
    run_.evens = evens;
    run_.ranges = {{start, limit}};
    // the edge values are asked about more than once: each Miller-Rabin
    // test is made once
    std::pair<uint64_t, bool> memo[8];
    size_t nmemo = 0;
    const auto prime = [&](uint64_t q) {
        for (size_t i = 0; i < nmemo; ++i)
            if (memo[i].first == q) return memo[i].second;
        const bool r = is_prime_u64(q);
        if (nmemo < 8) memo[nmemo++] = {q, r};
        return r;
    };
    // the device counted primes q in [start - 3, limit - 3]: the run's are
    // [start + 1, limit - 1]
    for (uint64_t q : {start - 3, start - 1})
        if (prime(q)) run_.primes[prime_slot(q)] -= 1;
    if (prime(limit - 1)) run_.primes[prime_slot(limit - 1)] += 1;
    // and twin pairs with the lower prime in [start - 5, limit - 5]: the
    // run's have it in [start + 1, limit - 3]
    for (uint64_t q : {start - 5, start - 3, start - 1})
        if (prime(q) && prime(q + 2)) run_.twins[twin_slot(q)] -= 1;
    if (prime(limit - 3) && prime(limit - 1)) run_.twins[twin_slot(limit - 3)] += 1;
    for (uint64_t& v : run_.primes)
        if (int64_t(v) < 0) throw std::runtime_error("stats: negative prime count after the edge corrections");
    if (gaps_from_ != start) return;
    GapBook& b = gaps_;
    if (last_prime_ == 0) last_prime_ = before_start_ = prev_prime(start - 3);   // no segment held a prime
    b.count = run_primes();
    if (limit < start + 2) return;
    // start - 1 lies outside the run: the walked twin gap start - 3 -> start - 1 goes
    if (prime(start - 3) && prime(start - 1)) b.walk(2, -1);
    // limit - 1 lies inside it, beyond the device's counting range
    if (prime(limit - 1)) {
        const uint64_t g = limit - 1 - last_prime_;
        b.exact(g, last_prime_);
        if (last_walked_) b.walk(g, +1);
    }
    // the sum of the gaps telescopes: last prime - the prime before the first
    uint64_t f = (start + 1) | 1;
    while (f <= limit - 1 && !prime(f)) f += 2;
    if (f <= limit - 1) {
        // the prime before f: nothing in (start, f) is prime, so it is
        // start - 1, start - 3 or prev_prime(start - 3), looked up above
        const uint64_t before_f = prime(start - 1) ? start - 1 : prime(start - 3) ? start - 3 : before_start_;
        const uint64_t last = prime(limit - 1) ? limit - 1 : prev_prime(limit - 1);
        b.span = last - before_f;
    }
}

uint64_t GoldbachStats::run_primes() const { return std::accumulate(std::begin(run_.primes), std::end(run_.primes), uint64_t(0)); }

uint64_t GoldbachStats::run_tail() const
{
    uint64_t t = run_.beyond.size();
    for (size_t cr = 0; cr < 15; ++cr)
        for (size_t i = tail0_; i < np_; ++i) t += run_.tail_cls[cr * np_ + i];
    for (size_t i = tail0_; i < np_; ++i) t += run_.tail_flat[i];
    return t;
}

uint64_t GoldbachStats::run_pmax() const
{
    uint64_t top = 0;
    for (const auto& [n, p] : run_.beyond) top = std::max(top, p);
    if (top) return top;
    for (size_t i = np_; i-- > tail0_;) {
        uint64_t c = run_.tail_flat[i];
        for (size_t cr = 0; cr < 15 && !c; ++cr) c = run_.tail_cls[cr * np_ + i];
        if (c) return cands_[i];
    }
    return 0;
}

void GoldbachStats::note_first(uint32_t pi, uint64_t n)
{
    uint64_t& f = run_.first[pi];
    const bool new_to_run = f == 0;
    if (new_to_run || n < f) f = n;
    if (new_to_run && camp_.first[pi] == 0 && sp_unknown(cands_[pi]))
        announce(n, cands_[pi], "SP",
                 "S(p) candidate: first even in this campaign with p(n) = p; S(p) was unknown (> 4e18) -- this IS "
                 "S(p) if the campaign covers [4e18, n] without a gap");
}

void GoldbachStats::announce(uint64_t n, uint64_t p, const char* kind, const char* what)
{
    // re-proved by an independent ascending scan before anything is written
    uint64_t witness = 0;
    for (uint64_t c = 3; c <= n / 2 && !witness; c += 2)
        if (is_prime_u64(c) && is_prime_u64(n - c)) witness = c;
    const char* verdict = witness == p ? "OK" : "MISMATCH";
    char stamp[32];
    const std::time_t now = std::time(nullptr);
    std::tm local{};
    localtime_r(&now, &local);
    std::strftime(stamp, sizeof stamp, "%Y-%m-%dT%H:%M:%S", &local);
    const std::string line = std::string(stamp) + " p(n)=" + std::to_string(p) + " n=" + std::to_string(n) +
                             " q=" + std::to_string(n - p) + " cpu_check=" + verdict + "  " + what + "\n";
    std::fprintf(stderr, "\n*** %s", line.c_str());
    if (!textio::append_file(rec_path_, line)) std::fprintf(stderr, "stats: cannot append to %s\n", rec_path_.c_str());
    events_.push_back({n, p});
    // the notification hook: GB_RECORD_HOOK, else the campaign's own script
    // when the working directory has one (run detached, output dropped)
    std::string hook;
    if (const char* h = std::getenv("GB_RECORD_HOOK")) hook = h;
    else if (textio::read_file("scripts/notify_record.sh")) hook = "bash scripts/notify_record.sh";
    if (!hook.empty()) {
        const std::string cmd = hook + " " + kind + " " + std::to_string(p) + " " + std::to_string(n) + " " +
                                std::to_string(n - p) + " " + verdict + " >/dev/null 2>&1 &";
        if (std::system(cmd.c_str()) != 0) std::fprintf(stderr, "record hook failed to start: %s\n", hook.c_str());
    }
}

std::vector<uint64_t> GoldbachStats::per_prime(const std::vector<uint64_t>& by_class,
                                               const std::vector<uint64_t>* flat) const
{
    std::vector<uint64_t> v = flat ? *flat : std::vector<uint64_t>(np_, 0);
    for (size_t cr = 0; cr < 15; ++cr)
        for (size_t i = 0; i < np_; ++i) v[i] += by_class[cr * np_ + i];
    return v;
}

std::string GoldbachStats::gap_fields() const
{
    using textio::append_dec;
    const GapBook& b = gaps_;
    std::string s;
    s.reserve(4096);
    s += " gaps=";
    append_dec(s, b.count);
    s += " gd1=";
    append_dec(s, b.span);
    s += " gmax=";
    if (b.top_exact() && b.top_g) {
        append_dec(s, b.top_g);
        s += ':';
        append_dec(s, b.top_lower);
    } else {
        s += '-';
    }
    s += " gtail=";
    if (b.tail.empty()) s += '-';
    bool more = false;
    for (const auto& [g, t] : b.tail) {
        if (more) s += ';';
        more = true;
        append_dec(s, uint64_t(g));
        s += ':';
        append_dec(s, t.count);
        s += ':';
        append_dec(s, t.first_lower);
    }
    s += " gwalk=";
    append_dec(s, walked_evens_);
    s += ':';
    append_dec(s, b.walk_count);
    s += " gsmom=";
    for (int k = 0; k < 4; ++k) {
        if (k) s += ',';
        append_dec(s, b.moment[k]);
    }
    s += " gsamp=";
    const size_t at = s.size();
    textio::append_count_list(s, b.walk_hist.data(), b.walk_hist.size());
    if (s.size() == at) s += '-';
    return s;
}

std::string GoldbachStats::log_line(const char* version) const
{
    using textio::append_dec;
    std::string line;
    line.reserve(8192);
    const auto num = [&](const char* k, uint64_t v) {
        line += ' ';
        line += k;
        line += '=';
        append_dec(line, v);
    };
    const auto csv = [&](const char* k, const uint64_t* v, size_t n) {
        line += ' ';
        line += k;
        line += '=';
        for (size_t i = 0; i < n; ++i) {
            if (i) line += ',';
            append_dec(line, v[i]);
        }
    };
    line += "GBSTATS v2 version=";
    line += version;
    num("start", start_);
    num("limit", limit_);
    num("evens", run_.evens);
    num("evens_sampled", run_.sampled);
    num("tail_t", tail_t_);
    num("tail_evens", run_tail());
    num("pmax", run_pmax());
    num("primes", run_primes());
    num("twins", run_twins());
    csv("primes30", run_.primes, 8);
    csv("twins30", run_.twins, 3);
    line += " new_first=";
    size_t at = line.size();
    for (size_t i = 0; i < np_; ++i) {
        const uint64_t f = run_.first[i], c = camp_.first[i];
        if (f && (!c || f < c)) {
            if (line.size() != at) line += ';';
            append_dec(line, uint64_t(cands_[i]));
            line += ':';
            append_dec(line, f);
        }
    }
    if (line.size() == at) line += '-';
    line += " beyond=";
    at = line.size();
    for (const auto& [n, p] : run_.beyond) {
        if (line.size() != at) line += ';';
        append_dec(line, p);
        line += ':';
        append_dec(line, n);
    }
    if (line.size() == at) line += '-';
    if (gaps_from_ == start_) line += gap_fields();
    line += " tail=";
    at = line.size();
    if (tail0_ < np_) {
        std::vector<uint64_t> t(run_.tail_flat.begin() + tail0_, run_.tail_flat.end());
        for (size_t cr = 0; cr < 15; ++cr)
            for (size_t i = tail0_; i < np_; ++i) t[i - tail0_] += run_.tail_cls[cr * np_ + i];
        append_dec(line, uint64_t(cands_[tail0_]));
        line += ':';
        const size_t lst = line.size();
        textio::append_count_list(line, t.data(), t.size());
        if (line.size() == lst) line.resize(at);   // an empty list: no "p:" head
    }
    if (line.size() == at) line += '-';
    line += '\n';
    return line;
}

void GoldbachStats::log_append(const std::string& line) const
{
    if (!textio::append_file(log_path_, line)) std::fprintf(stderr, "stats: cannot append to %s\n", log_path_.c_str());
}

bool GoldbachStats::log_has_gap_row(uint64_t a, uint64_t b) const
{
    const std::optional<std::string> all = textio::read_file(log_path_);
    if (!all) return false;
    size_t at = 0;
    while (at < all->size()) {
        size_t nl = all->find('\n', at);
        if (nl == std::string::npos) nl = all->size();
        const std::string_view line(all->data() + at, nl - at);
        at = nl + 1;
        // rows with gap fields: GBSTATS lines of gap runs and GBGAPS rows
        if (line.substr(0, 2) != "GB" || line.find(" gd1=") == std::string_view::npos) continue;
        uint64_t s = 0, l = 0;
        bool has_s = false, has_l = false;
        for (std::string_view w : textio::words(line)) {
            if (w.substr(0, 6) == "start=") has_s = textio::parse(w.substr(6), s);
            else if (w.substr(0, 6) == "limit=") has_l = textio::parse(w.substr(6), l);
        }
        if (has_s && has_l && s <= b && a <= l) return true;
    }
    return false;
}

// ---- the state file v2 ------------------------------------------------------

std::string GoldbachStats::state_text(const Tally& t) const
{
    using textio::append_count_list;
    using textio::append_dec;
    std::string s;
    s.reserve(size_t(16) << 10);
    s += "# goldbach-l0 statistics state v2 (exact; include/gstats.hpp)\n"
         "# one record per line; count lists are indexed by odd prime (3, 5, 7, ...), an empty entry is 0\n"
         "# sampled <n mod 30 / 2> <list>: p(n) of every even of the sampled segments\n"
         "# tailcls <n mod 30 / 2> <list>, tailagg <list>: p(n) >= tail_t (sampled evens by class, the rest)\n"
         "# first <p> <n>: the campaign's first even with p(n) = p, for p >= tail_t\n";
    s += "tail_t ";
    append_dec(s, uint64_t(tail_t_));
    s += '\n';
    for (const auto& [a, b] : t.ranges) {
        s += "range ";
        append_dec(s, a);
        s += ' ';
        append_dec(s, b);
        s += '\n';
    }
    s += "evens ";
    append_dec(s, t.evens);
    s += ' ';
    append_dec(s, t.sampled);
    s += "\nprimes30";
    for (uint64_t v : t.primes) {
        s += ' ';
        append_dec(s, v);
    }
    s += "\ntwins30";
    for (uint64_t v : t.twins) {
        s += ' ';
        append_dec(s, v);
    }
    s += '\n';
    for (size_t i = 0; i < np_; ++i)
        if (t.first[i]) {
            s += "first ";
            append_dec(s, uint64_t(cands_[i]));
            s += ' ';
            append_dec(s, t.first[i]);
            s += '\n';
        }
    for (const auto& [n, p] : t.beyond) {
        s += "beyond ";
        append_dec(s, p);
        s += ' ';
        append_dec(s, n);
        s += '\n';
    }
    for (const char* key : {"sampled ", "tailcls "}) {
        const std::vector<uint64_t>& v = key[0] == 's' ? t.hist : t.tail_cls;
        for (size_t cr = 0; cr < 15; ++cr) {
            s += key;
            append_dec(s, uint64_t(cr));
            s += ' ';
            append_count_list(s, v.data() + cr * np_, np_);
            s += '\n';
        }
    }
    s += "tailagg ";
    append_count_list(s, t.tail_flat.data(), t.tail_flat.size());
    s += '\n';
    return s;
}

void GoldbachStats::read_state(Tally& t) const
{
    t.clear();
    const std::optional<std::string> text = textio::read_file(state_path_);
    if (!text) return;   // a fresh campaign
    t.zero = false;      // its counts come from the file
    t.size_by_class(np_);
    const std::string what = "state " + state_path_;
    if (text->empty()) throw std::runtime_error("stats: " + what + " is empty (truncated write?); refusing to merge into it");
    const auto u64 = [](std::string_view w, uint64_t& v) { return textio::parse(w, v); };
    const std::vector<Rule> rules = {
        {"tail_t", 2, 2, 1, [&](const Words& w) {
             uint32_t v = 0;
             if (!textio::parse(w[1], v)) return false;
             if (v != tail_t_)
                 throw std::runtime_error("stats: " + what + " was recorded with tail_t " + std::to_string(v) +
                                          ", this run uses " + std::to_string(tail_t_));
             return true;
         }},
        {"range", 3, 3, 0, [&](const Words& w) {
             uint64_t a = 0, b = 0;
             if (!u64(w[1], a) || !u64(w[2], b) || a > b) return false;
             t.ranges.push_back({a, b});
             return true;
         }},
        {"evens", 3, 3, 1, [&](const Words& w) { return u64(w[1], t.evens) && u64(w[2], t.sampled); }},
        {"primes30", 9, 9, 1, [&](const Words& w) {
             for (size_t i = 0; i < 8; ++i)
                 if (!u64(w[1 + i], t.primes[i])) return false;
             return true;
         }},
        {"twins30", 4, 4, 1, [&](const Words& w) {
             for (size_t i = 0; i < 3; ++i)
                 if (!u64(w[1 + i], t.twins[i])) return false;
             return true;
         }},
        {"first", 3, 3, 0, [&](const Words& w) {
             uint32_t p = 0;
             uint64_t n = 0;
             if (!textio::parse(w[1], p) || !u64(w[2], n) || !(p & 1) || p > cands_.back() || index_of_[p / 2] < 0)
                 return false;
             t.first[size_t(index_of_[p / 2])] = n;
             return true;
         }},
        {"beyond", 3, 3, 0, [&](const Words& w) {
             uint64_t p = 0, n = 0;
             if (!u64(w[1], p) || !u64(w[2], n)) return false;
             t.beyond.push_back({n, p});
             return true;
         }},
        {"sampled", 3, 3, 15, [&](const Words& w) {
             uint32_t cr = 0;
             return textio::parse(w[1], cr) && textio::parse_count_list(w[2], t.hist.data() + cr * np_, np_);
         }},
        {"tailcls", 3, 3, 15, [&](const Words& w) {
             uint32_t cr = 0;
             return textio::parse(w[1], cr) && textio::parse_count_list(w[2], t.tail_cls.data() + cr * np_, np_);
         }},
        {"tailagg", 2, 2, 1, [&](const Words& w) { return textio::parse_count_list(w[1], t.tail_flat.data(), np_); }},
    };
    read_by_rules(*text, what, rules);
    normalize(t.ranges);
}

// ---- the gap state file v1 ---------------------------------------------------

GoldbachStats::GapFile GoldbachStats::read_gap_file() const
{
    GapFile f;
    const std::string path = state_path_ + ".gaps";
    const std::optional<std::string> text = textio::read_file(path);
    if (!text) return f;
    GapBook& b = f.book;
    const auto u64 = [](std::string_view w, uint64_t& v) { return textio::parse(w, v); };
    const std::vector<Rule> rules = {
        {"gaps_range", 3, 3, 0, [&](const Words& w) {
             uint64_t a = 0, c = 0;
             if (!u64(w[1], a) || !u64(w[2], c) || a > c) return false;
             f.ranges.push_back({a, c});
             return true;
         }},
        {"gaps_evens", 3, 3, 1, [&](const Words& w) { return u64(w[1], f.evens) && u64(w[2], f.walked); }},
        {"gaps", 3, 3, 1, [&](const Words& w) { return u64(w[1], b.count) && textio::parse(w[2], b.span); }},
        {"gaps_walk", 6, 6, 1, [&](const Words& w) {
             if (!u64(w[1], b.walk_count)) return false;
             for (size_t k = 0; k < 4; ++k)
                 if (!textio::parse(w[2 + k], b.moment[k])) return false;
             return true;
         }},
        {"gaps_samp", 1, 2, 1, [&](const Words& w) {
             return textio::parse_count_list(w.size() == 2 ? w[1] : std::string_view(), b.walk_hist.data(), GAP_BINS);
         }},
        {"gaps_max", 4, 4, 1, [&](const Words& w) {
             if (!u64(w[1], b.top_g) || !u64(w[2], b.top_lower) || (w[3] != "0" && w[3] != "1")) return false;
             b.all_walked = w[3] == "1" && b.top_g < GAP_TAIL;
             return true;
         }},
        {"gaps_tail", 4, 4, 0, [&](const Words& w) {
             uint32_t g = 0;
             uint64_t n = 0, lower = 0;
             if (!textio::parse(w[1], g) || !u64(w[2], n) || !u64(w[3], lower)) return false;
             b.add_tail(g, n, lower);
             return true;
         }},
        {"end", 1, 1, 1, [](const Words&) { return true; }},
    };
    read_by_rules(*text, "gap state " + path, rules);
    if (f.ranges.empty())
        throw std::runtime_error("stats: gap state " + path + " lacks its 'gaps_range' records (cut off or damaged); "
                                 "refusing to merge into it");
    normalize(f.ranges);
    return f;
}

bool GoldbachStats::commit_gap_file()
{
    GapFile f = read_gap_file();
    const std::string path = state_path_ + ".gaps";
    if (intersects(f.ranges, start_, limit_)) {
        std::fprintf(stderr, "stats: gaps of [%llu, %llu] already in %s -- not recorded again\n",
                     (unsigned long long)start_, (unsigned long long)limit_, path.c_str());
        return false;
    }
    f.book.absorb(gaps_);
    f.evens += run_.evens;
    f.walked += walked_evens_;
    f.ranges.push_back({start_, limit_});
    normalize(f.ranges);
    const GapBook& b = f.book;
    using textio::append_dec;
    std::string s;
    s.reserve(8192);
    s += "# goldbach-l0 prime gap state v1 (exact; include/gstats.hpp GapStats)\n"
         "# gaps_range <lo> <hi>: prime gaps recorded for the upper prime in [lo + 1, hi - 1]\n"
         "# gaps_evens <evens> <walked evens>; gaps <count> <sum of g> (every gap, exact)\n"
         "# gaps_walk <count> <sums of g, g^2, g^3, g^4>, gaps_samp <histogram by g / 2>: the walked segments\n"
         "# gaps_max <g> <lower> <1: the largest, 0: the largest seen>; gaps_tail <g> <count> <first lower>: g >= 700\n";
    for (const auto& [a, c] : f.ranges) {
        s += "gaps_range ";
        append_dec(s, a);
        s += ' ';
        append_dec(s, c);
        s += '\n';
    }
    s += "gaps_evens ";
    append_dec(s, f.evens);
    s += ' ';
    append_dec(s, f.walked);
    s += "\ngaps ";
    append_dec(s, b.count);
    s += ' ';
    append_dec(s, b.span);
    s += "\ngaps_walk ";
    append_dec(s, b.walk_count);
    for (const u128& m : b.moment) {
        s += ' ';
        append_dec(s, m);
    }
    s += "\ngaps_samp ";
    textio::append_count_list(s, b.walk_hist.data(), b.walk_hist.size());
    s += "\ngaps_max ";
    append_dec(s, b.top_g);
    s += ' ';
    append_dec(s, b.top_lower);
    s += b.top_exact() ? " 1\n" : " 0\n";
    for (const auto& [g, t] : b.tail) {
        s += "gaps_tail ";
        append_dec(s, uint64_t(g));
        s += ' ';
        append_dec(s, t.count);
        s += ' ';
        append_dec(s, t.first_lower);
        s += '\n';
    }
    s += "end\n";
    textio::replace_file(path, s);
    return true;
}

// ---- commit -------------------------------------------------------------------

std::string GoldbachStats::memory_commit(uint64_t sampled_total)
{
    if (campaign_) throw std::runtime_error("stats: memory_commit on a campaign run");
    if (run_.sampled != sampled_total)
        throw std::runtime_error("stats: " + std::to_string(run_.sampled) + " sampled evens recorded, " +
                                 std::to_string(sampled_total) + " in the sampled segments");
    camp_.clear();          // a fresh campaign that then holds this run alone
    camp_.absorb(run_);
    return state_text(camp_);
}

void GoldbachStats::commit(const char* version, uint64_t sampled_total, const std::string* run_line)
{
    if (!campaign_) throw std::runtime_error("stats: commit on a run without campaign files");
    if (run_.sampled != sampled_total)
        throw std::runtime_error("stats: " + std::to_string(run_.sampled) + " sampled evens recorded, " +
                                 std::to_string(sampled_total) + " in the sampled segments (statistics not written)");
    const textio::FileLock lock(state_path_ + ".lock");
    // re-read under the lock: another process may have committed meanwhile;
    // the caller's log line names first occurrences new to the campaign, so
    // it holds only while those are unchanged
    const std::vector<uint64_t> firsts_seen = camp_.first;
    read_state(camp_);
    const std::string line = run_line && camp_.first == firsts_seen ? *run_line : log_line(version);
    std::pair<uint64_t, uint64_t> covered;
    if (intersects(camp_.ranges, start_, limit_, &covered)) {
        std::fprintf(stderr, "stats: [%llu, %llu] overlaps the covered range [%llu, %llu] of %s -- not recorded again\n",
                     (unsigned long long)start_, (unsigned long long)limit_, (unsigned long long)covered.first,
                     (unsigned long long)covered.second, state_path_.c_str());
        if (gaps_from_ != start_) return;
        // what the campaign may still lack of an overlapping gap run: the
        // run's sampled histogram by prime, and its gap row
        const std::string head = " version=" + std::string(version) + " start=" + std::to_string(start_) +
                                 " limit=" + std::to_string(limit_) + " evens=" + std::to_string(run_.evens);
        log_append("GBPSAMP v1" + head + " psev=" + std::to_string(run_.sampled) +
                   " psamp=" + count_list(per_prime(run_.hist, nullptr)) + "\n");
        const std::string gap_row = "GBGAPS v1" + head + gap_fields() + "\n";
        if (commit_gap_file()) {
            log_append(gap_row);
        } else {
            const GapFile f = read_gap_file();
            bool inside = false;
            for (const auto& [a, b] : f.ranges) inside = inside || (a <= start_ && limit_ <= b);
            if (inside && !log_has_gap_row(start_, limit_)) {
                std::fprintf(stderr, "stats: %s has no gap row for [%llu, %llu] -- written now\n", log_path_.c_str(),
                             (unsigned long long)start_, (unsigned long long)limit_);
                log_append(gap_row);
            }
        }
        return;
    }
    camp_.absorb(run_);
    if (gaps_from_ == start_) commit_gap_file();
    textio::replace_file(state_path_, state_text(camp_));
    log_append(line);
}

void GoldbachStats::write_live(uint64_t start, uint64_t limit, uint64_t done, uint64_t total, double elapsed, bool force)
{
    if (!campaign_) return;
    const long long now = (long long)std::time(nullptr);
    if (!force && now - live_stamp_ < 2) return;
    live_stamp_ = now;
    // the run's largest p(n) so far and its first even; evens at or above 8000, 9000 and beyond 9781
    uint64_t pm = 0, pm_n = 0, at_least[3] = {};
    for (size_t i = tail0_; i < np_; ++i)
        if (run_.first[i] && cands_[i] > pm) {
            pm = cands_[i];
            pm_n = run_.first[i];
        }
    for (const auto& [n, p] : run_.beyond)
        if (p > pm) {
            pm = p;
            pm_n = n;
        }
    // per prime, its evens with p(n) >= tail_t (zero below tail0_, where
    // nothing is counted as tail)
    const uint64_t bounds[3] = {8000, 9000, RECORD_P + 1};
    uint64_t tail_evens = run_.beyond.size();
    for (size_t i = tail0_; i < np_; ++i) {
        uint64_t t = run_.tail_flat[i];
        for (size_t cr = 0; cr < 15; ++cr) t += run_.tail_cls[cr * np_ + i];
        tail_evens += t;
        for (int k = 0; k < 3; ++k)
            if (cands_[i] >= bounds[k]) at_least[k] += t;
    }
    at_least[2] += run_.beyond.size();
    char head[512];
    std::snprintf(head, sizeof head,
                  "{\"time\": %lld, \"start\": %llu, \"limit\": %llu, \"evens_done\": %llu, \"evens_total\": %llu, "
                  "\"elapsed_s\": %.1f, \"tail_t\": %u, \"tail_evens\": %llu, \"pmax\": %llu, \"pmax_n\": %llu, "
                  "\"ge8000\": %llu, \"ge9000\": %llu, \"gt9781\": %llu, \"primes\": %llu, \"twins\": %llu, "
                  "\"events\": [",
                  now, (unsigned long long)start, (unsigned long long)limit, (unsigned long long)done,
                  (unsigned long long)total, elapsed, tail_t_, (unsigned long long)tail_evens, (unsigned long long)pm,
                  (unsigned long long)pm_n, (unsigned long long)at_least[0], (unsigned long long)at_least[1],
                  (unsigned long long)at_least[2], (unsigned long long)run_primes(), (unsigned long long)run_twins());
    std::string json = head;
    for (size_t i = 0; i < events_.size(); ++i)
        json += std::string(i ? ", " : "") + "[" + std::to_string(events_[i].second) + ", \"" +
                std::to_string(events_[i].first) + "\"]";
    json += "]}\n";
    try {   // advisory: replaced atomically but not synced; a failed write leaves the last one
        textio::replace_file(live_path_, json, ".tmp", false);
    } catch (const std::exception&) {
    }
}

}  // namespace gb
