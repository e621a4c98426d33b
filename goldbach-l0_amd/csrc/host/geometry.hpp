#pragma once
// Segment and class-map geometry of the dense path, and the per-segment
// device argument blocks derived from it.
//
// Segmentation and the class layout follow the reference exactly (its
// statistics depend on them: sampling is per segment index, counting ranges
// per segment): src/host_math.cpp:13-43 segment_geometry,
// src/host_math.cpp:63-78 class_geometry, include/gstats.hpp:262-296
// count30_limits. The device blocks (window size, strides, mask residues)
// are this build's own.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "config.hpp"
#include "gbh.h"

namespace gb {

inline constexpr uint32_t kRes[8] = {1, 7, 11, 13, 17, 19, 23, 29};
inline constexpr uint32_t kSmallP[15] = {7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};

struct SegmentGeometry {
    uint64_t seg_start = 0, seg_end = 0, seg_evens = 0;
    uint64_t q_first = 0, q_last = 0, q_odds = 0;
};

/// One segment of a run: [seg_start, seg_end] with the short remainder
/// merged into the last segment; q covers [seg_start - 9781, seg_end - 3].
/// @throws std::runtime_error when the segment has no odd q (n = 4 alone).
inline SegmentGeometry segment_geometry(uint64_t seg_start, uint64_t limit, uint64_t span)
{
    SegmentGeometry g;
    g.seg_start = seg_start;
    g.seg_end = (limit - seg_start <= span) ? limit : seg_start + span - 2;
    g.seg_evens = (g.seg_end - g.seg_start) / 2 + 1;
    g.q_first = seg_start >= MAX_GB_PRIME + 2 ? seg_start - MAX_GB_PRIME : 3;
    g.q_last = g.seg_end - 3;
    if (g.q_last < g.q_first)
        throw std::runtime_error("segment [" + std::to_string(seg_start) + ", " + std::to_string(g.seg_end) +
                                 "] has no odd q to sieve");
    g.q_odds = (g.q_last - g.q_first) / 2 + 1;
    return g;
}

struct ClassGeo {
    uint32_t phi[8] = {};   // class array c starts at q_first + 2 phi[c]
    uint64_t Q[8] = {};
    uint64_t n[8] = {};     // bits of class array c
};

inline ClassGeo class_geometry(uint64_t q_first, uint64_t q_odds)
{
    ClassGeo g;
    const uint32_t r0 = uint32_t(q_first % 30);
    for (int c = 0; c < 8; ++c) {
        g.phi[c] = ((kRes[c] + 30 - r0) % 30) / 2;
        g.Q[c] = q_first + 2ull * g.phi[c];
        g.n[c] = q_odds > g.phi[c] ? (q_odds - g.phi[c] + 14) / 15 : 0;
    }
    return g;
}

inline uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t m) { return uint64_t((unsigned __int128)a * b % m); }

inline uint64_t inv_mod(uint64_t a, uint64_t m)   // m prime, a coprime
{
    uint64_t r = 1, e = m - 2, b = a % m;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod64(r, b, m);
        b = mulmod64(b, b, m);
    }
    return r;
}

/// Device geometry block of a segment. @p wbits is the LDS window in bits.
inline gbh_geom make_geom(const SegmentGeometry& sg, const ClassGeo& cg, uint32_t wbits)
{
    gbh_geom g{};
    g.seg_start = sg.seg_start;
    g.seg_end = sg.seg_end;
    g.q_first = sg.q_first;
    g.q_last = sg.q_last;
    uint64_t nmax = 1;
    for (int c = 0; c < 8; ++c) {
        g.Q[c] = cg.Q[c];
        if (cg.n[c] >= (1ull << 31)) throw std::runtime_error("segment class array exceeds 2^31 bits");
        g.nc[c] = uint32_t(cg.n[c]);
        nmax = std::max<uint64_t>(nmax, cg.n[c]);
    }
    g.nwin = uint32_t((nmax + wbits - 1) / wbits);
    g.np_win = 0xFFFFFFFFu;   // every window prime (callers bound it by the range's sqrt)
    g.ridge_t = 0xFFFFFFFFu;  // no ridge records
    for (int c = 0; c < 8; ++c) {
        uint64_t restore = 0;
        for (int k = 0; k < 15; ++k) {
            const uint64_t p = kSmallP[k];
            // class bit t of every multiple of p: Q + 30 t = 0 (mod p)
            const uint64_t qm = cg.Q[c] % p;
            g.s0[c][k] = uint32_t(mulmod64((p - qm) % p, inv_mod(30 % p, p), p));
            // p itself lies in the class array: keep it prime
            if (p % 30 == kRes[c] && p >= cg.Q[c] && p <= sg.q_last) {
                const uint64_t t = (p - cg.Q[c]) / 30;
                if (t < 64) restore |= 1ull << t;
            }
        }
        g.restore[c] = restore;
    }
    return g;
}

/// count30_limits: class c counts bits t with Q[c] + 30 t in [lo, hi].
inline void class_range(const ClassGeo& cg, int c, uint64_t lo, uint64_t hi, uint32_t& tl, uint32_t& th)
{
    tl = lo <= cg.Q[c] ? 0u : uint32_t((lo - cg.Q[c] + 29) / 30);
    th = hi < cg.Q[c] ? 0u : uint32_t((hi - cg.Q[c]) / 30 + 1);
    if (th < tl) th = tl;
}

/// Counter arguments of a segment: primes q in [seg_start - 3, seg_end - 3],
/// twin pairs (q, q + 2) with q in [seg_start - 5, seg_end - 5] (the ranges
/// of consecutive segments tile the q axis).
/// @pre seg_start >= 6.
inline gbh_countseg make_count_args(const SegmentGeometry& sg, const ClassGeo& cg)
{
    gbh_countseg a{};
    a.on = 1;
    int order[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    std::sort(order, order + 8, [&](int x, int y) { return cg.phi[x] < cg.phi[y]; });
    uint32_t nbits = 0;
    for (int k = 0; k < 8; ++k) {
        const int c = order[k];
        a.cls[k] = uint32_t(c);
        a.pos0[k] = 2 * cg.phi[c];
        class_range(cg, c, sg.seg_start - 3, sg.seg_end - 3, a.lo[k], a.hi[k]);
        nbits = std::max(nbits, a.hi[k]);
    }
    static constexpr int kPair[3][2] = {{2, 3}, {4, 5}, {7, 0}};   // 11/13, 17/19, 29/31
    for (int j = 0; j < 3; ++j) {
        const int k = kPair[j][0], k2 = kPair[j][1];
        const int64_t d = int64_t(cg.Q[k]) + 2 - int64_t(cg.Q[k2]);
        if (d != 0 && d != 30) throw std::runtime_error("count: twin class offset");
        a.tk[j] = uint32_t(k);
        a.tk2[j] = uint32_t(k2);
        a.td[j] = uint32_t(d / 30);
        class_range(cg, k, sg.seg_start - 5, sg.seg_end - 5, a.tlo[j], a.thi[j]);
        nbits = std::max(nbits, a.thi[j]);
    }
    a.nwords = (nbits + 63) / 64;
    return a;
}

}  // namespace gb
