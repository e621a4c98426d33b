// Draws the seeded inputs of the throughput slice and the batched workload
// and prints them (the committed lists cfg3_segments.txt and cfg4_ranges.txt
// are its output):
//   cfg3: 16 segment indices of the 1e12-integer slice at 4.5e18 (2140
//         segments of 467,500,000 integers, the last one short), seed 0,
//         Floyd's sampling without replacement, ascending;
//   cfg4: 256 ranges, seed 1: start = 2 floor(exp(U(ln 1e9, ln 5e18)) / 2),
//         length = floor(exp(U(ln 1e5, ln 4.675e8))) integers,
//         limit = start + length rounded down to even, at most 2^64 - 2^40.
// Both use std::mt19937_64 with libstdc++'s distributions.
//     g++ -O2 -std=c++17 -o /tmp/draw_configs tests/golden/draw_configs.cpp
//     /tmp/draw_configs cfg3 > tests/golden/cfg3_segments.txt
//     /tmp/draw_configs cfg4 > tests/golden/cfg4_ranges.txt
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <set>

int main(int argc, char** argv)
{
    if (argc == 2 && !std::strcmp(argv[1], "cfg3")) {
        const uint64_t start = 4500000000000000000ull, limit = start + 1000000000000ull - 2, span = 467500000ull;
        const uint64_t nseg = (limit - start) / span + 1;   // the short remainder is its own last segment
        std::mt19937_64 g(0);
        std::set<uint64_t> s;
        for (uint64_t j = nseg - 16; j < nseg; ++j) {
            const uint64_t t = std::uniform_int_distribution<uint64_t>(0, j)(g);
            s.insert(s.count(t) ? j : t);
        }
        std::printf("# segment index, first even, last even (slice [%llu, %llu], span %llu, %llu segments)\n",
                    (unsigned long long)start, (unsigned long long)limit, (unsigned long long)span,
                    (unsigned long long)nseg);
        for (uint64_t k : s) {
            const uint64_t a = start + k * span;
            const uint64_t b = (limit - a <= span) ? limit : a + span - 2;
            std::printf("%llu %llu %llu\n", (unsigned long long)k, (unsigned long long)a, (unsigned long long)b);
        }
        return 0;
    }
    if (argc == 2 && !std::strcmp(argv[1], "cfg4")) {
        std::mt19937_64 g(1);
        std::uniform_real_distribution<double> us(std::log(1e9), std::log(5e18)), ul(std::log(1e5), std::log(4.675e8));
        const uint64_t top = ~0ull - (1ull << 40) + 1;      // 2^64 - 2^40
        std::printf("# start (first even), limit (last even), length drawn; 256 ranges, seed 1\n");
        for (int i = 0; i < 256; ++i) {
            const uint64_t start = 2 * uint64_t(std::floor(std::exp(us(g)) / 2));
            const uint64_t len = uint64_t(std::floor(std::exp(ul(g))));
            uint64_t limit = (start + len) & ~1ull;
            if (limit > top) limit = top;
            std::printf("%llu %llu %llu\n", (unsigned long long)start, (unsigned long long)limit,
                        (unsigned long long)len);
        }
        return 0;
    }
    std::fprintf(stderr, "usage: draw_configs cfg3|cfg4\n");
    return 2;
}
