// Host unit tests (no GPU) of the verifier's host layer: the text codecs and
// file helpers, Miller-Rabin and the base-prime sieve, the GBP2 prime cache,
// the ridge board and store, and the statistics' campaign files (state v2,
// gap state v1, log rows, overlap refusal, damaged files). Built plain and
// with AddressSanitizer + UBSan (goldbach-l0_amd/Makefile, target
// host_unit) and run by tests/test_host_unit.py. Exit status 0 when every
// check holds; each failed check is printed.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gbh.h"
#include "gstats.hpp"
#include "primes.hpp"
#include "ridge.hpp"
#include "textio.hpp"

namespace {

int g_failed = 0, g_checks = 0;

#define CHECK(cond)                                                                      \
    do {                                                                                 \
        ++g_checks;                                                                      \
        if (!(cond)) {                                                                   \
            ++g_failed;                                                                  \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
        }                                                                                \
    } while (0)

template <class F>
bool throws(F&& f)
{
    try {
        f();
    } catch (const std::exception&) {
        return true;
    }
    return false;
}

std::string g_dir;

std::string path(const std::string& name) { return g_dir + "/" + name; }
std::string slurp(const std::string& p) { return gb::textio::read_file(p).value_or("<absent>"); }
bool exists(const std::string& p)
{
    struct stat st{};
    return ::stat(p.c_str(), &st) == 0;
}
std::vector<std::string> lines_with(const std::string& text, const std::string& head)
{
    std::vector<std::string> out;
    size_t at = 0;
    while (at < text.size()) {
        size_t nl = text.find('\n', at);
        if (nl == std::string::npos) nl = text.size();
        const std::string l = text.substr(at, nl - at);
        if (l.compare(0, head.size(), head) == 0) out.push_back(l);
        at = nl + 1;
    }
    return out;
}

// ---------------------------------------------------------------- textio
void test_textio()
{
    using namespace gb::textio;
    const uint64_t v[] = {0, 5, 0, 3, 0, 0};
    CHECK(count_list(v, 6) == ",5,,3");
    CHECK(count_list(v, 1).empty());
    uint64_t back[6];
    CHECK(parse_count_list(",5,,3", back, 6) && std::equal(back, back + 6, v));
    CHECK(parse_count_list("", back, 6) && back[1] == 0);
    CHECK(!parse_count_list("1,2,3", back, 2));     // more entries than room
    CHECK(!parse_count_list("1,x", back, 6));
    uint64_t u = 0;
    CHECK(parse("18446744073709551615", u) && u == UINT64_MAX);
    CHECK(!parse("18446744073709551616", u));
    CHECK(!parse("", u) && !parse("1a", u) && !parse("-1", u) && !parse(" 1", u));
    uint32_t w = 0;
    CHECK(parse("4294967295", w) && w == 4294967295u && !parse("4294967296", w));
    const u128 top = ~u128(0);
    CHECK(dec(top) == "340282366920938463463374607431768211455");
    u128 t = 0;
    CHECK(parse(dec(top), t) && t == top && !parse("340282366920938463463374607431768211456", t));
    CHECK(dec(0) == "0");
    const auto ws = words("a  b ");
    CHECK(ws.size() == 4 && ws[0] == "a" && ws[1].empty() && ws[2] == "b" && ws[3].empty());
    replace_file(path("f.txt"), "one\n");
    CHECK(slurp(path("f.txt")) == "one\n" && !exists(path("f.txt.tmp")));
    CHECK(append_file(path("f.txt"), "two\n") && slurp(path("f.txt")) == "one\ntwo\n");
    CHECK(!read_file(path("absent")).has_value());
    CHECK(throws([] { replace_file(path("no/such/dir/f"), "x"); }));
}

// ---------------------------------------------------------------- primes
void test_primes()
{
    using gb::is_prime_u64;
    CHECK(is_prime_u64(2) && is_prime_u64(3) && !is_prime_u64(1) && !is_prime_u64(0) && !is_prime_u64(9));
    CHECK(is_prime_u64((uint64_t(1) << 61) - 1));
    CHECK(is_prime_u64(18446744073709551557ull));      // the largest prime below 2^64
    CHECK(!is_prime_u64(3825123056546413051ull));      // a strong pseudoprime to bases 2..23
    CHECK(!is_prime_u64(uint64_t(4294967291ull) * 4294967279ull));
    CHECK(gb::prev_prime(100) == 97 && gb::prev_prime(4) == 3);
    const auto small = gb::odd_primes_up_to(100);
    CHECK(small.size() == 24 && small.front() == 3 && small.back() == 97);
    CHECK(gb::base_primes(2000003, 4) == gb::odd_primes_up_to(2000003));
    const uint32_t c[] = {3, 5, 7, 11};
    CHECK(gb::cpu_min_goldbach_prime(4, c, 4, 11) == 2 && gb::cpu_min_goldbach_prime(98, c, 4, 11) == 19);
}

// ---------------------------------------------------------------- packed large primes
void test_packed_primes()
{
    // every prime of a packed stream decodes back, across group edges,
    // including the largest prime gap below 2^32 (336 after 3842610773)
    std::vector<uint32_t> p;
    for (uint64_t x = 3842600001ull; x < 3842620000ull; x += 2)
        if (gb::is_prime_u64(x)) p.push_back(uint32_t(x));
    const auto it = std::find(p.begin(), p.end(), 3842610773u);
    CHECK(it != p.end() && it + 1 != p.end() && it[1] - it[0] == 336);
    for (uint32_t group : {64u, 16u}) {
        const gb::PackedPrimes pk = gb::pack_primes(p.data(), uint32_t(p.size()), group, 4096);
        CHECK(pk.n == p.size() && pk.gap.size() == pk.base.size() * group);
        CHECK(pk.base.size() == (p.size() + 4096 + group - 1) / group);
        bool same = true;
        for (uint32_t j = 0; j < p.size(); ++j) same = same && gb::packed_prime_at(pk, j, group) == p[j];
        CHECK(same);
        for (uint32_t j = uint32_t(p.size()); j < pk.gap.size(); ++j) same = same && pk.gap[j] == 0;
        CHECK(same);
        // past the last prime: bases hold the last prime, so the base after
        // any group bounds every prime of it
        CHECK(pk.base.back() == p.back() && pk.base[(p.size() - 1) / group + 1] >= p.back());
        for (uint32_t g = 0; g < p.size() / group; ++g) same = same && pk.gap[g * group] == 0 && pk.base[g] == p[g * group];
        CHECK(same);
    }
    // a half gap of 255 fits, 256 does not; gaps must be even and ascending
    const uint32_t fit[] = {1000003, 1000003 + 510}, wide[] = {1000003, 1000003 + 512}, odd[] = {5, 8};
    CHECK(gb::packed_prime_at(gb::pack_primes(fit, 2, 64, 0), 1, 64) == 1000003 + 510);
    CHECK(throws([&] { gb::pack_primes(wide, 2, 64, 0); }));
    CHECK(throws([&] { gb::pack_primes(odd, 2, 64, 0); }));
    CHECK(throws([&] { gb::pack_primes(fit, 2, 48, 0); }));
}

// ---------------------------------------------------------------- prime cache
void test_prime_cache()
{
    const std::string stem = path("pc");
    const uint32_t n = 100000;
    const auto want = gb::odd_primes_up_to(n);
    CHECK(gb::base_primes_cached(n, stem.c_str(), 2) == want);
    const std::string f = stem + "." + std::to_string(n);
    const std::string raw = slurp(f);
    CHECK(raw.size() == 28 + 4 * want.size() && raw.compare(0, 4, "GBP2") == 0);
    uint64_t hdr[3];
    std::memcpy(hdr, raw.data() + 4, 24);
    CHECK(hdr[0] == n && hdr[1] == want.size() && hdr[2] == gb::prime_payload_hash(want.data(), want.size()));
    CHECK(gb::base_primes_cached(n, stem.c_str(), 2) == want);   // read back
    std::string bad = raw;
    bad[100] ^= 1;
    gb::textio::replace_file(f, bad);
    CHECK(gb::base_primes_cached(n, stem.c_str(), 2) == want);   // ignored, rebuilt, published again
    CHECK(slurp(f) == raw);
    gb::textio::replace_file(f, raw.substr(0, raw.size() - 4));  // cut off
    CHECK(gb::base_primes_cached(n, stem.c_str(), 2) == want && slurp(f) == raw);
    // a smaller bound: the prefix of the larger file, nothing new written
    CHECK(gb::base_primes_cached(5000, stem.c_str(), 2) == gb::odd_primes_up_to(5000));
    CHECK(!exists(stem + ".5000"));
}

// ---------------------------------------------------------------- ridges
uint64_t smallest_p(uint64_t n)
{
    for (uint64_t p = 3;; p += 2)
        if (gb::is_prime_u64(p) && gb::is_prime_u64(n - p)) return p;
}

void test_ridges()
{
    ::setenv("GB_RIDGE_MIN", "3", 1);
    const auto cand = gb::odd_primes_up_to(9781);
    gb::RidgeTracker r(&cand);
    CHECK(r.active() && r.floor() == 3 && r.thresh() == 3);
    std::vector<gb::RidgeEntry> all;
    for (uint64_t n = 100000; n < 101000; n += 2) {
        const uint64_t p = smallest_p(n);
        all.push_back({n, p, n - p, 99000, 1});
        r.consider(n, p, 99000, 1);
    }
    std::sort(all.begin(), all.end(), [](auto& a, auto& b) { return a.p > b.p || (a.p == b.p && a.n > b.n); });
    all.resize(30);
    CHECK(r.top().size() == 30);
    bool same = true;
    for (size_t i = 0; i < 30 && i < r.top().size(); ++i) same = same && r.top()[i].n == all[i].n && r.top()[i].p == all[i].p;
    CHECK(same);
    CHECK(r.thresh() == all.back().p);
    const uint64_t dropped = r.dropped();
    r.consider(100002, 101, 99000);     // 100002 - 101 is not prime: dropped
    r.consider(all[0].n + 1000002, 1000003, 99000);   // p not minimal (or not prime): dropped or below
    CHECK(r.dropped() >= dropped + 1);
    // device records with an overflow: the stored ones count, the run is flagged
    const uint64_t pairs[4] = {all[0].n, all[0].p, all[1].n, all[1].p};
    r.merge(pairs, 2, 5, 99000);
    CHECK(r.overflowed());
    r.save(path("ridges.csv"));
    const std::string csv = slurp(path("ridges.csv"));
    CHECK(csv.rfind("# goldbach-l0 ridge store v2", 0) == 0);
    CHECK(lines_with(csv, "1,").size() == 1 && lines_with(csv, "1," + std::to_string(all[0].n)).size() == 1);
    gb::RidgeTracker back(&cand);
    CHECK(back.load(path("ridges.csv")) && back.top().size() == 30);
    bool kept = true;
    for (size_t i = 0; i < 30 && i < back.top().size(); ++i)
        kept = kept && back.top()[i].n == r.top()[i].n && back.top()[i].found == r.top()[i].found;
    CHECK(kept);
    // v1 rows (no found time) and junk rows
    gb::textio::replace_file(path("r1.csv"), "# v1\n1," + std::to_string(all[0].n) + "," + std::to_string(all[0].p) + "," +
                                                 std::to_string(all[0].q) + ",99000\nnot,a,row\n");
    gb::RidgeTracker v1(&cand);
    CHECK(v1.load(path("r1.csv")) && v1.top().size() == 1 && v1.top()[0].found == 0);
    ::unsetenv("GB_RIDGE_MIN");
}

// ---------------------------------------------------------------- statistics
std::vector<uint16_t> g_slot_pi[15];

void feed(gb::GoldbachStats& s, uint64_t start, uint64_t limit, uint64_t seed, uint64_t* sampled)
{
    // one segment's worth of synthetic device outputs: prime / twin counts
    // (large enough for the edge corrections), a sampled histogram and the
    // tail's first occurrences
    uint32_t cnt[8], tw[3];
    for (int i = 0; i < 8; ++i) cnt[i] = uint32_t(5000 + 37 * i + seed);
    for (int i = 0; i < 3; ++i) tw[i] = uint32_t(300 + i + seed);
    s.add_counts(cnt, tw);
    std::vector<uint32_t> hist(15 * GBH_NCAND, 0);
    for (uint32_t cr = 0; cr < 15; ++cr)
        for (uint32_t k = 0; k < 40; ++k) hist[cr * GBH_NCAND + k] = (cr + k + uint32_t(seed)) % 7;
    hist[3 * GBH_NCAND + 500] = 2;   // a tail prime (index 500: p > 3000)
    *sampled += s.add_hist_slots(true, hist.data(), GBH_NCAND, g_slot_pi);
    std::vector<uint32_t> first(1205, ~0u);
    first[500] = uint32_t(seed + 10);
    s.add_first(start, first.data());
    s.finish_run(start, limit, (limit - start) / 2 + 1);
}

void test_stats()
{
    const auto cand = gb::odd_primes_up_to(9781);
    for (uint32_t cr = 0; cr < 15; ++cr) {
        g_slot_pi[cr].resize(1205);
        for (uint32_t k = 0; k < 1205; ++k) g_slot_pi[cr][k] = uint16_t(k);
    }
    ::setenv("GB_STATS_STATE", path("st.state").c_str(), 1);
    ::setenv("GB_STATS_LOG", path("st.log").c_str(), 1);
    ::setenv("GB_RECORDS_FILE", path("rec.txt").c_str(), 1);
    ::setenv("GB_STATS_LIVE", path("live.json").c_str(), 1);
    ::setenv("GB_RECORD_HOOK", "", 1);
    std::string first_state, first_line;
    {
        gb::GoldbachStats s(cand, 3000, true);
        s.set_run_start(1000000);
        uint64_t sampled = 0;
        feed(s, 1000000, 2000000, 1, &sampled);
        first_line = s.log_line("t");
        s.commit("t", sampled);
        first_state = slurp(path("st.state"));
        CHECK(lines_with(first_state, "range ") == std::vector<std::string>{"range 1000000 2000000"});
        CHECK(lines_with(first_state, "evens ") ==
              std::vector<std::string>{"evens 500001 " + std::to_string(sampled)});
        CHECK(lines_with(first_state, "sampled ").size() == 15 && lines_with(first_state, "tailcls ").size() == 15);
        CHECK(lines_with(first_state, "first " + std::to_string(cand[500]) + " 1000022").size() == 1);
        CHECK(lines_with(slurp(path("st.log")), "GBSTATS v2 version=t start=1000000 limit=2000000 evens=500001").size() == 1);
        CHECK(first_line.find(" new_first=" + std::to_string(cand[500]) + ":1000022 ") != std::string::npos);
        CHECK(first_line.find(" tail=3001:") != std::string::npos);
        CHECK(s.run_pmax() == cand[500]);
        // the same run without campaign files holds the same state
        gb::GoldbachStats m(cand, 3000, false);
        m.set_run_start(1000000);
        uint64_t ms = 0;
        feed(m, 1000000, 2000000, 1, &ms);
        CHECK(m.log_line("t") == first_line);
        CHECK(m.memory_commit(ms) == first_state);
    }
    {   // the same range again: refused, nothing changes
        gb::GoldbachStats s(cand, 3000, true);
        s.set_run_start(1000000);
        uint64_t sampled = 0;
        feed(s, 1000000, 2000000, 1, &sampled);
        const std::string log_before = slurp(path("st.log"));
        s.commit("t", sampled);
        CHECK(slurp(path("st.state")) == first_state && slurp(path("st.log")) == log_before);
    }
    {   // an adjacent range merges; its first occurrence is not new to the campaign
        gb::GoldbachStats s(cand, 3000, true);
        s.set_run_start(2000002);
        CHECK(!s.wanted(500));
        uint64_t sampled = 0;
        feed(s, 2000002, 3000000, 2, &sampled);
        CHECK(s.log_line("t").find(" new_first=- ") != std::string::npos);
        s.commit("t", sampled);
        const std::string st = slurp(path("st.state"));
        CHECK(lines_with(st, "range ") == std::vector<std::string>{"range 1000000 3000000"});
        CHECK(lines_with(st, "evens 1000001 ").size() == 1);
        CHECK(throws([&] { s.commit("t", sampled + 1); }));   // a sampled-count mismatch is refused
    }
    {   // damaged and cut-off state files fail before any merge
        const std::string good = slurp(path("st.state"));
        gb::textio::replace_file(path("st.state"), good + "bogus 1\n");
        CHECK(throws([&] { gb::GoldbachStats s(cand, 3000, true); }));
        gb::textio::replace_file(path("st.state"), good.substr(0, good.size() - 3));
        CHECK(throws([&] { gb::GoldbachStats s(cand, 3000, true); }));
        std::string twice = good;
        twice += "tailagg 1\n";
        gb::textio::replace_file(path("st.state"), twice);
        CHECK(throws([&] { gb::GoldbachStats s(cand, 3000, true); }));
        std::string other_t = good;
        other_t.replace(other_t.find("tail_t 3000"), 11, "tail_t 2000");
        gb::textio::replace_file(path("st.state"), other_t);
        CHECK(throws([&] { gb::GoldbachStats s(cand, 3000, true); }));
        gb::textio::replace_file(path("st.state"), good);
        CHECK(!throws([&] { gb::GoldbachStats s(cand, 3000, true); }));
    }
}

void test_gaps()
{
    const auto cand = gb::odd_primes_up_to(9781);
    ::setenv("GB_STATS_STATE", path("g.state").c_str(), 1);
    ::setenv("GB_STATS_LOG", path("g.log").c_str(), 1);
    const uint64_t start = 1000000, limit = 1000200;
    const auto run = [&](bool campaign, std::string* line) {
        gb::GoldbachStats s(cand, 3000, campaign);
        s.set_run_start(start);
        s.begin_gaps(start);
        // the primes of [start - 3, limit - 3] as a walked segment (the
        // device's positions are q - q_first)
        const uint64_t q_first = start - 3;
        std::vector<uint64_t> pr;
        for (uint64_t q = q_first; q <= limit - 3; q += 2)
            if (gb::is_prime_u64(q)) pr.push_back(q);
        std::vector<uint32_t> hist(gb::GAP_BINS + 1, 0);
        uint64_t gmax = 0;
        for (size_t i = 1; i < pr.size(); ++i) {
            const uint64_t g = pr[i] - pr[i - 1];
            ++hist[g / 2];
            gmax = std::max(gmax, (g << 32) | uint32_t(~uint32_t(pr[i] - q_first)));
        }
        gb::SegmentGaps sg;
        sg.q_first = q_first;
        sg.evens = (limit - start) / 2 + 1;
        sg.primes = pr.size();
        sg.pfirst = uint32_t(pr.front() - q_first);
        sg.plast = uint32_t(pr.back() - q_first);
        sg.walked = true;
        sg.gmax = gmax;
        sg.hist = hist.data();
        s.add_gap_segment(sg);
        uint32_t cnt[8] = {}, tw[3] = {};
        for (uint64_t q : pr) {
            static const int slot[30] = {-1, 0, -1, -1, -1, -1, -1, 1, -1, -1, -1, 2, -1, 3, -1,
                                         -1, -1, 4, -1, 5, -1, -1, -1, 6, -1, -1, -1, -1, -1, 7};
            ++cnt[slot[q % 30]];
        }
        s.add_counts(cnt, tw);
        s.finish_run(start, limit, (limit - start) / 2 + 1);
        *line = s.log_line("t");
        if (campaign) s.commit("t", 0);
        return s.run_gaps();
    };
    std::string line;
    const gb::GapBook b = run(true, &line);
    // the run's primes q in [start + 1, limit - 1] and their gaps from the
    // prime before each
    std::vector<uint64_t> pr;
    for (uint64_t q = start + 1; q <= limit - 1; q += 2)
        if (gb::is_prime_u64(q)) pr.push_back(q);
    uint64_t walk = 0, top = 0, top_lo = 0;
    gb::u128 m1 = 0;
    for (uint64_t q : pr) {
        const uint64_t g = q - gb::prev_prime(q);
        ++walk;
        m1 += g;
        if (g > top) { top = g; top_lo = q - g; }
    }
    CHECK(b.count == pr.size() && b.walk_count == walk && b.moment[0] == m1 && b.span == m1);
    CHECK(b.top_g == top && b.top_lower == top_lo && b.top_exact());
    CHECK(line.find(" gaps=" + std::to_string(pr.size()) + " gd1=" + gb::u128_str(m1) + " gmax=" +
                    std::to_string(top) + ":" + std::to_string(top_lo) + " ") != std::string::npos);
    const std::string gs = slurp(path("g.state.gaps"));
    CHECK(lines_with(gs, "gaps_range ") == std::vector<std::string>{"gaps_range 1000000 1000200"});
    CHECK(lines_with(gs, "gaps " + std::to_string(pr.size()) + " ").size() == 1 && lines_with(gs, "end").size() == 1);
    // again over the same range: only the sampled-histogram row is logged
    const std::string log_before = slurp(path("g.log"));
    run(true, &line);
    const std::string log_after = slurp(path("g.log"));
    CHECK(log_after.size() > log_before.size() && lines_with(log_after, "GBPSAMP v1 ").size() == 1 &&
          lines_with(log_after, "GBGAPS v1 ").empty() && slurp(path("g.state.gaps")) == gs);
    // a damaged gap state fails before the run
    gb::textio::replace_file(path("g.state.gaps"), gs.substr(0, gs.find("end")));
    CHECK(throws([&] { gb::GoldbachStats s(cand, 3000, true); }));
}

}  // namespace

int main()
{
    char tmpl[] = "/tmp/gb_host_unit.XXXXXX";
    if (!::mkdtemp(tmpl)) return 2;
    g_dir = tmpl;
    test_textio();
    test_primes();
    test_packed_primes();
    test_prime_cache();
    test_ridges();
    test_stats();
    test_gaps();
    std::printf("host_unit: %d checks, %d failed\n", g_checks, g_failed);
    std::string rm = "rm -rf '" + g_dir + "'";
    if (std::system(rm.c_str()) != 0) std::fprintf(stderr, "could not remove %s\n", g_dir.c_str());
    return g_failed ? 1 : 0;
}
