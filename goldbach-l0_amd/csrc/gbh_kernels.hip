// Device kernels of the per-segment verification pipeline, written for
// gfx950 (CDNA4: wave64, 160 KB LDS per CU). All arithmetic is integer.
//
// Map layout (same contract as the reference's class map, goldbach.cl:234-252
// and host_math.hpp:88-107): eight class arrays per segment, one per residue
// r in {1,7,11,13,17,19,23,29} mod 30; bit t of class array c is 1 iff
// q = Q[c] + 30 t is prime (for q in [q_first, q_last]). Words are 64-bit,
// little-endian, so the same memory read as 32-bit words gives bit t at
// word t >> 5, bit t & 31 (the LDS atomics use that view). A batch of
// consecutive segments keeps its maps side by side: map + (s * 8 + c) * cw.
//
// Stages (one launch each per batch, in stream order):
//   k_setup     first class bit of every window prime, per segment and class
//               (reference mid_setup, goldbach.cl:3528-3563)
//   k_harvest   hit lists of the large primes over the whole batch, filed by
//               (segment, class, window): multiples p m with m prime from
//               the multiplier bitmap (reference hl_harvest_mt, 3251-3426)
//   k_mtrans    builds the multiplier bitmap from class maps sieved over m
//               (reference mt_build, 3181-3242)
//   k_window    one work-group per (window, class, segment): masks for
//               7..61, LDS marks of the window primes, the window's hit list,
//               one coalesced store (reference make_prime_map + mid_window +
//               hl_apply, 1092-1608, 3600-3701, 3434-3457)
//   k_coff      per-segment candidate table of the checker
//   k_check_a   exact ascending candidates, 64 evens per lane, first ka
//               candidates (reference check30_a, 2028-2189)
//   k_check_b   the rest for every pending even (reference check30_b,
//               2210-2306)
//   k_count_a/b primes, twins and the gap walk of the counting range
//               (reference count30 + gap_tail_exact, 2499-2825)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gbh.h"
#include "gbh_tables.h"

namespace gbh_dev {

__constant__ uint16_t c_cand_p[GBH_EVEN_CLASSES * GBH_NCAND];   // useful candidates ascending, per even class
__constant__ uint8_t c_cand_c[GBH_EVEN_CLASSES * GBH_NCAND];    // class array of q = n - p
__constant__ uint16_t c_cand_pi[GBH_EVEN_CLASSES * GBH_NCAND];  // prime index of the candidate
__constant__ uint32_t c_cand_n[GBH_EVEN_CLASSES];               // useful candidates per even class

__device__ __forceinline__ uint32_t wave_lane() { return __lane_id(); }

// Exclusive prefix over the wave of v (64 lanes); returns the total in *tot.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t* tot)
{
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if ((int)wave_lane() >= d) x += y;
    }
    *tot = __shfl(x, 63, 64);
    return x - v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Inclusive prefix sum over the wave's 64 lanes (every lane active): DPP
// shifts by 1, 2, 4 and 8 inside each row of 16 lanes, then the row totals
// carried by row_bcast:15 (into rows 1 and 3) and row_bcast:31 (into rows 2
// and 3). Two 16-bit halves scan independently while no half's total
// reaches 2^16.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);    // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);    // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);    // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);    // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return v;
}

// The packed large-prime stream (gbh_pstream): the primes of whole groups,
// one per lane. jA, jB: the first index of a group each (multiples of 64,
// uniform over the wave, which must be whole). A group's half-gaps sum
// below 63 * 256 < 2^16, so both groups scan in one word.
// Exclusive prefix over the wave (every lane active) of v; the total in *tot.
__device__ __forceinline__ uint32_t wave_excl_scan_dpp(uint32_t v, uint32_t* tot)
{
    const uint32_t x = wave_incl_scan(v);
    *tot = (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
    return x - v;
}

// In two steps, so the loads can be in flight while other work is done:
// ps_load2 issues them (the lane's two gaps in one word, the two bases),
// ps_finish2 scans.
struct PsPair {
    uint32_t gaps, bA, bB;
};
// The group bases through the constant address space: the stream is never
// written while a kernel reads it, so a uniform index is a scalar load.
typedef const __attribute__((address_space(4))) uint32_t* ConstU32;
__device__ __forceinline__ uint32_t ps_base(const gbh_pstream& ps, uint32_t g)
{
    return ((ConstU32)ps.base)[g];
}
__device__ __forceinline__ PsPair ps_load2(const gbh_pstream& ps, uint32_t jA, uint32_t jB, uint32_t lane)
{
    const uint32_t gA = __builtin_amdgcn_readfirstlane(jA >> 6), gB = __builtin_amdgcn_readfirstlane(jB >> 6);
    return PsPair{(uint32_t)ps.gap[(size_t)gA * 64 + lane] | ((uint32_t)ps.gap[(size_t)gB * 64 + lane] << 16),
                  ps_base(ps, gA), ps_base(ps, gB)};
}
__device__ __forceinline__ void ps_finish2(const PsPair& x, uint32_t& pA, uint32_t& pB)
{
    const uint32_t v = wave_incl_scan(x.gaps);
    pA = x.bA + 2 * (v & 0xFFFFu);
    pB = x.bB + 2 * (v >> 16);
}
__device__ __forceinline__ void ps_decode2(const gbh_pstream& ps, uint32_t jA, uint32_t jB, uint32_t lane,
                                           uint32_t& pA, uint32_t& pB)
{
    ps_finish2(ps_load2(ps, jA, jB, lane), pA, pB);
}

__device__ __forceinline__ uint32_t ps_decode1(const gbh_pstream& ps, uint32_t jA, uint32_t lane)
{
    const uint32_t gA = __builtin_amdgcn_readfirstlane(jA >> 6);
    return ps_base(ps, gA) + 2 * wave_incl_scan((uint32_t)ps.gap[(size_t)gA * 64 + lane]);
}

__device__ __forceinline__ uint64_t* seg_map(const gbh_batch& b, uint32_t s)
{
    return b.map + (size_t)s * 8 * b.cw;
}

// The smallest multiplier m >= max(ceil(q_first / p), p): multiples are
// marked from max(p^2, q_first), so a prime p inside the range stays prime.
__device__ __forceinline__ uint64_t first_mult(uint64_t q_first, uint32_t p)
{
    uint64_t m0 = (q_first + p - 1) / p;
    if (m0 < p) m0 = p;
    return m0;
}

// ---------------------------------------------------------------- stage 1
__global__ void __launch_bounds__(256) k_setup(gbh_batch b, const uint32_t* __restrict__ primes, uint32_t np,
                                               uint32_t* __restrict__ i0)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = blockIdx.y;
    const gbh_geom& g = b.geoms[s];
    if (j >= np || j >= g.np_win) return;   // primes above sqrt(q_last) have no multiple to mark (k_window skips them)
    const uint32_t p = primes[j];
    const uint64_t m0 = first_mult(g.q_first, p);
    const uint32_t m0r = (uint32_t)(m0 % 30);
    const uint32_t pinv = kInv30[p % 30];
    uint32_t* out = i0 + (size_t)s * 8 * np;
#pragma unroll
    for (int c = 0; c < GBH_CLASSES; ++c) {
        const uint32_t mt = (kRes[c] * pinv) % 30;
        const uint64_t m = m0 + (mt + 30 - m0r) % 30;
        const uint64_t q = m * p;
        uint32_t t = 0xFFFFFFFFu;
        if (q <= g.q_last && q >= g.Q[c]) {
            const uint64_t tt = (q - g.Q[c]) / 30;
            if (tt < g.nc[c]) t = (uint32_t)tt;
        }
        out[(size_t)c * np + j] = t;
    }
}

// ---------------------------------------------------------------- stage 2
// Large primes p >= P (the window bound, with P^3 above every q of the
// batch) are listed per batch, filed by (segment, class, window) so the
// window sieve applies each window's list in LDS.
//
// A composite q with no prime factor below P has exactly two prime factors
// (q < P^3), so q = p m with p <= m and m prime. Primes p >= t_hl therefore
// list only the multiples p m whose multiplier m is prime, read from the
// resident multiplier bitmap (reference mt_build / hl_harvest_mt,
// goldbach.cl:3181-3426; here the bitmap covers every m the run can need
// instead of a capped table). Primes in [P, t_hl) (present only when the
// bitmap budget caps m) walk every multiplier coprime to 30030 (reference
// sieve_large_packed's wheel, goldbach.cl:480-748).
//
// Multiplier bitmap: word w, bit 8 j + i <-> m = m_base + 240 w + 30 j + R[i]
// with R the unit residues mod 30 ascending, so the bits ascend with m.
//
// Positions inside a batch are half-offsets d = (q - qb0) / 2 < 2^32 (the
// host keeps a batch's q range below 2^33), so the per-hit work is 32-bit.

constexpr uint64_t kResPacked = 0x1D1713110D0B0701ull;   // R[i] in byte i
constexpr uint32_t kUnitMask = 0x208A2882u;               // bit r set for the units r mod 30

__device__ __forceinline__ uint32_t res_of(uint32_t i) { return (uint32_t)(kResPacked >> (8 * i)) & 0xFFu; }
// unit residues mod 30 below r, r in [0, 30]
__device__ __forceinline__ uint32_t units_below(uint32_t r) { return (uint32_t)__popc(kUnitMask & ((1u << r) - 1)); }

struct HlCtx {
    uint64_t qb0;            // the batch's first q
    uint32_t dmax;           // half-offset of the batch's last q
    uint32_t hs_magic;       // floor(2^32 / half span): estimates d / half span from below
    uint32_t nseg, nwin, lw, wmask;
    uint32_t t_hl;           // first prime served by the bitmap
    const uint64_t* mbits;
    uint64_t m_base, m_span; // the bitmap covers m in [m_base, m_base + m_span)
    uint32_t* stv;           // staged entries of the counting pass (nullptr: no staging)
    uint16_t* stb;           // their lists
    uint32_t* sn;            // entries offered to the stage
    uint32_t scap;
    uint32_t* region;        // kDirect: the work-group's region, [bucket][rcap]
    uint32_t rcap;
};

// Harvest passes: kCount counts per window list (and stages), kWrite writes
// the window lists, kDirect (one pass) appends to the work-group's own
// region of each (segment, class) list.
enum { kCount = 0, kWrite = 1, kDirect = 2 };

// Files the multiple at half-offset d of class c into the window list of
// every segment whose q range holds it (consecutive segments overlap by 9776
// in q); the entry is the multiple's bit inside its window (kDirect: its bit
// in the class array).
template <int kMode>
__device__ __forceinline__ void file_one(const HlCtx& h, uint32_t s, uint32_t c, uint32_t d, const uint32_t* dq,
                                         uint32_t* cnt, uint32_t* __restrict__ hits, uint32_t cap,
                                         uint32_t* __restrict__ err)
{
    const uint32_t t = __umulhi(d - dq[s * 8 + c], 0x88888889u) >> 3;   // (q - Q) / 30 = (d - dq) / 15
    if (kMode == kDirect) {
        const uint32_t pos = atomicAdd(&cnt[s * 8 + c], 1u);
        if (pos < h.rcap) h.region[(size_t)(s * 8 + c) * h.rcap + pos] = t;
        else atomicOr(err, 1u);
        return;
    }
    const uint32_t bk = (s * 8 + c) * h.nwin + (t >> h.lw);
    if (kMode == kCount) {
        atomicAdd(&cnt[bk], 1u);
        if (h.stv) {
            const uint32_t pos = atomicAdd(h.sn, 1u);
            if (pos < h.scap) {
                h.stv[pos] = t & h.wmask;
                h.stb[pos] = (uint16_t)bk;
            }
        }
    } else {
        const uint32_t pos = atomicAdd(&cnt[bk], 1u);
        if (pos < cap) hits[(size_t)bk * cap + pos] = t & h.wmask;
        else atomicOr(err, 1u);
    }
}

template <int kMode>
__device__ __forceinline__ void file_hit(const HlCtx& h, const uint32_t* df, const uint32_t* dl, const uint32_t* dq,
                                         uint32_t d, uint32_t c, uint32_t* cnt, uint32_t* __restrict__ hits,
                                         uint32_t cap, uint32_t* __restrict__ err)
{
    // the segment starting at or below d: an estimate at most two below,
    // corrected without branches
    uint32_t s = min(__umulhi(d, h.hs_magic), h.nseg - 1);
    s += (s + 1 < h.nseg && d >= df[min(s + 1, h.nseg - 1)]);
    s += (s + 1 < h.nseg && d >= df[min(s + 1, h.nseg - 1)]);
    while (s + 1 < h.nseg && d >= df[s + 1]) ++s;     // irregular spacing only (a run's first segment below 9783)
    while (s > 0 && d < df[s]) --s;
    if (d <= dl[s]) file_one<kMode>(h, s, c, d, dq, cnt, hits, cap, err);
    if (s > 0 && d <= dl[s - 1]) file_one<kMode>(h, s - 1, c, d, dq, cnt, hits, cap, err);   // overlap
}

// 1 / p in double: the hardware reciprocal and one Newton step
__device__ __forceinline__ double inv_of(uint32_t p)
{
    const double pd = (double)p;
    double r = __builtin_amdgcn_rcp(pd);
    const double e = __builtin_fma(-pd, r, 1.0);
    return __builtin_fma(r, e, r);
}

// floor(x / p) for x < 2^63, p >= 2^12 and its remainder: a double-precision
// quotient (x rounds to within 2^10, inv and the product to about 2^-50
// relative, so the quotient, below 2^51, is within one) made exact by one
// remainder step each way. Returns *rem >= p if the estimate was further off
// (the caller fails closed). Written without correction loops: a loop of
// r += p would be compiled into a 64-bit division.
__device__ __forceinline__ uint64_t div_big(uint64_t x, double xd, uint32_t p, double inv, uint32_t* rem)
{
    const double qd = xd * inv;
    const uint64_t hi = (uint64_t)(uint32_t)(qd * (1.0 / 4294967296.0));
    uint64_t q = (hi << 32) + (uint64_t)(uint32_t)(qd - (double)hi * 4294967296.0);
    int64_t r = (int64_t)(x - q * p);
    if (r < 0) { --q; r += p; }
    else if (r >= (int64_t)p) { ++q; r -= p; }
    *rem = r >= 0 && r < (int64_t)p ? (uint32_t)r : p;
    return q;
}

// The wheel walk: every multiple q = p m in [qb0, qb1] with m >= p and m
// coprime to 30030 (a multiple of 2 3 5 7 11 13 is not in the class map or
// is cleared by the mask stage): m walks the 48 residues coprime to 210, 11
// and 13 are tested by running residues, the class of q follows from m mod 30.
struct WheelLds {
    uint8_t next[210], idx[210], inc[48], cls[30];
};

template <int kMode>
__device__ __forceinline__ void walk_wheel(const HlCtx& h, double qb0d, uint32_t p, const uint32_t* df,
                                           const uint32_t* dl, const uint32_t* dq, const WheelLds& wl, uint32_t* cnt,
                                           uint32_t* __restrict__ hits, uint32_t cap, uint32_t* __restrict__ err)
{
    uint32_t rem;
    uint64_t m = div_big(h.qb0, qb0d, p, inv_of(p), &rem);
    if (rem >= p) { atomicOr(err, 16u); return; }
    m += rem != 0;
    if (m < p) m = p;
    const uint32_t r = (uint32_t)(m % 30030);
    const uint32_t r210 = r % 210;
    const uint32_t d0 = wl.next[r210];
    m += d0;                                   // m odd from here on: q - qb0 is even
    const uint64_t dd = (m * p - h.qb0) >> 1;
    if (dd > h.dmax) return;
    uint32_t d = (uint32_t)dd;
    uint32_t r11 = (r + d0) % 11, r13 = (r + d0) % 13, r30 = (r + d0) % 30;
    uint32_t idx = wl.idx[(r210 + d0) % 210];
    const uint32_t pr = p % 30;
    for (;;) {
        if (r11 && r13) file_hit<kMode>(h, df, dl, dq, d, wl.cls[(r30 * pr) % 30], cnt, hits, cap, err);
        const uint32_t inc = wl.inc[idx];       // even: the walk stays on odd m
        idx = idx == 47 ? 0 : idx + 1;
        const uint64_t step = (uint64_t)(inc >> 1) * p;
        if (step > h.dmax - d) break;
        d += (uint32_t)step;
        r11 += inc; if (r11 >= 11) r11 -= 11;
        r13 += inc; if (r13 >= 13) r13 -= 13;
        r30 += inc; if (r30 >= 30) r30 -= 30;
    }
}

// floor(x / 240) for x < 2^42: a double quotient, corrected once
__device__ __forceinline__ uint64_t words_of(uint64_t x)
{
    const double v = (double)x * (1.0 / 240.0);
    const uint32_t hi = (uint32_t)(v * (1.0 / 4294967296.0));
    uint64_t w = ((uint64_t)hi << 32) + (uint32_t)(v - (double)hi * 4294967296.0);
    if (w * 240 > x) --w;
    else if ((w + 1) * 240 <= x) ++w;
    return w;
}

// The bitmap walk: every multiple q = p m in [qb0, qb1] with m >= p prime,
// in two steps so a lane can have the first bitmap words of several primes
// in flight: bm_setup finds the word range, bm_run files the hits.
// cmt[k] holds, 3 bits per residue index i, the class of R[k] R[i] mod 30.
struct BmWalk {
    uint64_t w0, w1;              // first and last bitmap word
    uint64_t lo_mask, hi_mask;    // bits of m >= m0 in w0, of m <= m1 in w1
    uint32_t p, cmap;
    bool live;
};

__device__ __forceinline__ BmWalk bm_setup(const HlCtx& h, double qb0d, uint32_t p, const uint32_t* cmt,
                                           uint32_t* __restrict__ err)
{
    BmWalk k;
    k.live = false;
    k.p = p;
    k.w0 = k.w1 = 0;
    k.lo_mask = k.hi_mask = 0;
    k.cmap = 0;
    if (!p) return k;
    // m = floor(qb0 / p) from a double quotient (1 / p: the float reciprocal
    // and one Newton step, ~2^-44 relative), made exact by the remainder
    const double pd = (double)p;
    double inv = (double)__builtin_amdgcn_rcpf((float)p);
    inv = __builtin_fma(inv, __builtin_fma(-pd, inv, 1.0), inv);
    const double fd = __builtin_floor(qb0d * inv);
    const uint32_t fhi = (uint32_t)(fd * (1.0 / 4294967296.0));
    uint64_t m = ((uint64_t)fhi << 32) | (uint32_t)(fd - (double)fhi * 4294967296.0);
    // within one of floor(qb0 / p) (relative error below 2^-45 of a
    // quotient below 2^44): one step each way (no correction loop, which
    // would be compiled into a 64-bit division), then the guard
    int64_t r = (int64_t)(h.qb0 - m * p);
    if (r < 0) { --m; r += p; }
    else if (r >= (int64_t)p) { ++m; r -= p; }
    if (r < 0 || r >= (int64_t)p) { atomicOr(err, 16u); return k; }
    uint32_t r0 = 0;                       // m0 p = qb0 + r0
    if (r) { ++m; r0 = p - (uint32_t)r; }
    // further multipliers up to qb1: t = floor((span - r0) / p)
    const uint64_t span = 2ull * h.dmax;
    if (r0 > span) return k;
    const uint64_t y = span - r0;
    uint32_t t = 0;
    if (y >= p) {
        // y < 2^33 and p >= 2^12: the float quotient is within one
        t = (uint32_t)((float)(uint32_t)(y >> 1) * 2.0f * __builtin_amdgcn_rcpf((float)p));
        if ((uint64_t)t * p > y) --t;
        else if ((uint64_t)(t + 1) * p <= y) ++t;
        if ((uint64_t)t * p > y || (uint64_t)(t + 1) * p <= y) { atomicOr(err, 16u); return k; }
    } else if (m >= p) {
        // one multiplier only (the common case of a short batch): it must be
        // a unit mod 30 (2^32 = 16 mod 30) to be in the bitmap at all
        const uint32_t r1 = ((uint32_t)(m >> 32) * 16u + (uint32_t)m % 30u) % 30u;
        if (!((kUnitMask >> r1) & 1u)) return k;
    }
    uint64_t m0 = m;
    if (m0 < p) {                          // multipliers below p belong to smaller primes
        if ((uint64_t)p > m0 + t) return k;
        t -= (uint32_t)(p - m0);
        m0 = p;
    }
    if (m0 < h.m_base || m0 + t - h.m_base >= h.m_span) {   // outside the bitmap: fail closed
        atomicOr(err, 8u);
        return k;
    }
    // byte y0 = floor((m0 - m_base) / 30) and residue of m0 (m_base is a
    // multiple of 30; m0 - m_base < 2^42 is exact in double and its fused
    // quotient by 30 within 2^-15 of the true one, whose fractional parts
    // are multiples of 1/30: the floor of the quotient plus 2^-7 is exact),
    // then of m1 = m0 + t
    const uint64_t dm = m0 - h.m_base;
    const double dmd = __builtin_fma((double)(uint32_t)(dm >> 32), 4294967296.0, (double)(uint32_t)dm);
    const double fq = __builtin_floor(__builtin_fma(dmd, 1.0 / 30.0, 0x1p-7));
    const uint32_t fq_hi = (uint32_t)(fq * (1.0 / 4294967296.0));
    const uint64_t y0 = ((uint64_t)fq_hi << 32) | (uint32_t)(fq - (double)fq_hi * 4294967296.0);
    const uint32_t r30 = (uint32_t)dm - 30u * (uint32_t)y0;
    const uint32_t e = r30 + t;
    const uint64_t y1 = y0 + e / 30;
    const uint32_t r1 = e % 30;
    k.w0 = y0 >> 3;
    k.w1 = y1 >> 3;
    const uint32_t n0 = 8 * (uint32_t)(y0 & 7) + units_below(r30);     // bits of m < m0 in word w0 (<= 63)
    const uint32_t n1 = 8 * (uint32_t)(y1 & 7) + units_below(r1 + 1);  // bits of m <= m1 in word w1 (0..64)
    k.lo_mask = ~0ull << n0;
    k.hi_mask = n1 >= 64 ? ~0ull : ((1ull << n1) - 1);
    if (k.w0 == k.w1 && !(k.lo_mask & k.hi_mask)) return k;           // no multiplier coprime to 30
    k.cmap = cmt[units_below(p % 30)];
    k.live = true;
    return k;
}

// v: word w0 of the bitmap (already loaded). Returns whether any bit of the
// range was set.
template <int kMode>
__device__ __forceinline__ bool bm_run(const HlCtx& h, const BmWalk& k, uint64_t v, const uint32_t* df,
                                       const uint32_t* dl, const uint32_t* dq, uint32_t* cnt,
                                       uint32_t* __restrict__ hits, uint32_t cap, uint32_t* __restrict__ err)
{
    if (!k.live) return false;
    bool any = false;
    for (uint64_t w = k.w0;; ) {
        if (w == k.w0) v &= k.lo_mask;
        if (w == k.w1) v &= k.hi_mask;
        if (v) {
            any = true;
            const uint64_t qw = (uint64_t)k.p * (h.m_base + 240 * w) - h.qb0;   // mod 2^64; q - qb0 >= 0 for listed bits
            do {
                const uint32_t bit = (uint32_t)__ffsll((long long)v) - 1;
                v &= v - 1;
                const uint32_t i = bit & 7;
                const uint32_t delta = 30 * (bit >> 3) + res_of(i);
                const uint32_t d = (uint32_t)((qw + (uint64_t)k.p * delta) >> 1);
                file_hit<kMode>(h, df, dl, dq, d, (k.cmap >> (3 * i)) & 7, cnt, hits, cap, err);
            } while (v);
        }
        if (w >= k.w1) break;
        ++w;
        v = __builtin_nontemporal_load(h.mbits + w);
    }
    return any;
}

// kG bitmap primes per lane at a time: primes j0 + (k + u) 256 + tid for
// u < kG; on the writing pass only those flagged by the counting pass.
constexpr uint32_t kG = 4;
// bm_single's primes per lane at a time (its per-prime state is small)
constexpr uint32_t kS = 8;

template <int kMode>
__device__ __forceinline__ void bm_group(const HlCtx& h, double qb0d, const gbh_pstream& ps,
                                         uint32_t j0, uint32_t j1, uint32_t k, uint32_t tid, const uint32_t* df,
                                         const uint32_t* dl, const uint32_t* dq, const uint32_t* cmt, uint32_t* cnt,
                                         uint64_t* any_hit, uint32_t* __restrict__ hits, uint32_t cap,
                                         uint32_t* __restrict__ err)
{
    uint32_t p[kG];
    const uint32_t jw = j0 + (tid & ~63u);
#pragma unroll
    for (uint32_t u = 0; u < kG; u += 2)
        ps_decode2(ps, jw + (k + u) * 256, jw + (k + u + 1) * 256, tid & 63, p[u], p[u + 1]);
#pragma unroll
    for (uint32_t u = 0; u < kG; ++u) {
        const uint32_t j = j0 + (k + u) * 256 + tid;
        bool on = j < j1;
        if (kMode == kWrite && on) on = (any_hit[(k + u) * 4 + (tid >> 6)] >> (tid & 63)) & 1;
        p[u] = on ? p[u] : 0u;
    }
    BmWalk wk[kG];
    uint64_t v[kG];
#pragma unroll
    for (uint32_t u = 0; u < kG; ++u) wk[u] = bm_setup(h, qb0d, p[u], cmt, err);
#pragma unroll
    // bitmap words are read once: streamed past L2, whose lines then keep
    // the hit lists being written
    for (uint32_t u = 0; u < kG; ++u) v[u] = wk[u].live ? __builtin_nontemporal_load(h.mbits + wk[u].w0) : 0ull;
#pragma unroll
    for (uint32_t u = 0; u < kG; ++u) {
        const bool any = bm_run<kMode>(h, wk[u], v[u], df, dl, dq, cnt, hits, cap, err);
        if (kMode == kCount) {
            const uint64_t bal = __ballot(any);
            if ((tid & 63) == 0) any_hit[(k + u) * 4 + (tid >> 6)] = bal;
        }
    }
}

// Direct form, bitmap primes with several multipliers: the bitmap walk of
// bm_group with each wave's hits queued in LDS and filed by all its lanes
// at once. Filed where they are found, every lane of a wave waits for the
// lane with the most hits (their number varies from prime to prime) while
// its neighbours idle; queued, a word step costs one wave scan and the
// filing of ceil(hits / 64) hits per lane.
constexpr uint32_t kHitQ = 512;          // queued hits per wave at most
struct HitQLds {
    uint32_t d[4][kHitQ];                // half-offset of each queued hit
    uint8_t c[4][kHitQ];                 // its class
};

// This wave's LDS writes visible to its own later reads, in order.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Files the wave's n queued hits (every lane takes every 64th) and empties
// the queue; n is wave-uniform.
__device__ __forceinline__ void hitq_flush(const HlCtx& h, const uint32_t* df, const uint32_t* dl, const uint32_t* dq,
                                           uint32_t* cnt, uint32_t* __restrict__ err, const uint32_t* qd,
                                           const uint8_t* qc, uint32_t& n, uint32_t lane)
{
    wave_lds_sync();
    for (uint32_t e = lane; e < n; e += 64) file_hit<kDirect>(h, df, dl, dq, qd[e], qc[e], cnt, nullptr, 0, err);
    wave_lds_sync();
    n = 0;
}

// kGQ primes per lane at a time (primes j0 + (k + u) 256 + tid, u < kGQ);
// *n: the wave's queued hits (flushed by the caller at the end of the chunk).
constexpr uint32_t kGQ = 4;
__device__ __forceinline__ void bm_group_q(const HlCtx& h, double qb0d, const gbh_pstream& ps,
                                           uint32_t j0, uint32_t j1, uint32_t k, uint32_t tid, const uint32_t* df,
                                           const uint32_t* dl, const uint32_t* dq, const uint32_t* cmt, uint32_t* cnt,
                                           HitQLds& Q, uint32_t& n, uint32_t* __restrict__ err)
{
    const uint32_t lane = tid & 63, wv = tid >> 6;
    uint32_t* qd = Q.d[wv];
    uint8_t* qc = Q.c[wv];
    uint32_t p[kGQ];
    const uint32_t jw = j0 + (tid & ~63u);
#pragma unroll
    for (uint32_t u = 0; u < kGQ; u += 2)
        ps_decode2(ps, jw + (k + u) * 256, jw + (k + u + 1) * 256, lane, p[u], p[u + 1]);
#pragma unroll
    for (uint32_t u = 0; u < kGQ; ++u) {
        const uint32_t j = j0 + (k + u) * 256 + tid;
        p[u] = j < j1 ? p[u] : 0u;
    }
    BmWalk wk[kGQ];
    uint64_t v[kGQ];
#pragma unroll
    for (uint32_t u = 0; u < kGQ; ++u) wk[u] = bm_setup(h, qb0d, p[u], cmt, err);
#pragma unroll
    for (uint32_t u = 0; u < kGQ; ++u) v[u] = wk[u].live ? __builtin_nontemporal_load(h.mbits + wk[u].w0) : 0ull;
    for (uint32_t u = 0; u < kGQ; ++u) {
        const BmWalk& kk = wk[u];
        uint64_t w = kk.w0, vv = v[u];
        bool more = kk.live;
        for (;;) {                            // one bitmap word per lane and step; wave-uniform exit
            uint64_t x = 0;
            if (more) {
                x = vv;
                if (w == kk.w0) x &= kk.lo_mask;
                if (w == kk.w1) x &= kk.hi_mask;
            }
            uint32_t tot;
            const uint32_t off = wave_excl_scan_dpp((uint32_t)__popcll(x), &tot);
            if (tot) {
                if (n + tot > kHitQ) hitq_flush(h, df, dl, dq, cnt, err, qd, qc, n, lane);
                if (x) {
                    const uint64_t qw = (uint64_t)kk.p * (h.m_base + 240 * w) - h.qb0;   // mod 2^64
                    if (tot <= kHitQ) {
                        uint32_t at = n + off;
                        do {
                            const uint32_t bit = (uint32_t)__ffsll((long long)x) - 1;
                            x &= x - 1;
                            const uint32_t i = bit & 7;
                            const uint32_t delta = 30 * (bit >> 3) + res_of(i);
                            qd[at] = (uint32_t)((qw + (uint64_t)kk.p * delta) >> 1);
                            qc[at] = (uint8_t)((kk.cmap >> (3 * i)) & 7);
                            ++at;
                        } while (x);
                    } else {                  // more than a queue holds (dense words): filed in place
                        do {
                            const uint32_t bit = (uint32_t)__ffsll((long long)x) - 1;
                            x &= x - 1;
                            const uint32_t i = bit & 7;
                            const uint32_t delta = 30 * (bit >> 3) + res_of(i);
                            const uint32_t d = (uint32_t)((qw + (uint64_t)kk.p * delta) >> 1);
                            file_hit<kDirect>(h, df, dl, dq, d, (kk.cmap >> (3 * i)) & 7, cnt, nullptr, 0, err);
                        } while (x);
                    }
                }
                if (tot <= kHitQ) n += tot;
            }
            more = more && w < kk.w1;
            if (!__ballot(more)) break;
            if (more) {
                ++w;
                vv = __builtin_nontemporal_load(h.mbits + w);
            }
        }
    }
}

// Bitmap primes above the run's q span (p > qb1 - qb0) have at most one
// multiple in it, q = m0 p with m0 = ceil(qb0 / p): a lean form of
// bm_setup + bm_run for them (one quotient, one bit of one word; most
// primes leave after the span or unit test, before any load).
template <int kMode>
__device__ __forceinline__ void bm_single(const HlCtx& h, double qb0d, const gbh_pstream& ps,
                                          uint32_t j0, uint32_t j1, uint32_t k, uint32_t tid, const uint32_t* df,
                                          const uint32_t* dl, const uint32_t* dq, const uint32_t* cmt, uint32_t* cnt,
                                          uint64_t* any_hit, uint32_t* __restrict__ hits, uint32_t cap,
                                          uint32_t* __restrict__ err)
{
    uint32_t p[kS];
    const uint32_t jw = j0 + (tid & ~63u);
#pragma unroll
    for (uint32_t u = 0; u < kS; u += 2)
        ps_decode2(ps, jw + (k + u) * 256, jw + (k + u + 1) * 256, tid & 63, p[u], p[u + 1]);
#pragma unroll
    for (uint32_t u = 0; u < kS; ++u) {
        const uint32_t j = j0 + (k + u) * 256 + tid;
        bool on = j < j1;
        if (kMode == kWrite && on) on = (any_hit[(k + u) * 4 + (tid >> 6)] >> (tid & 63)) & 1;
        p[u] = on ? p[u] : 0u;
    }
    const uint64_t span = 2ull * h.dmax;
    uint64_t w[kS];
    uint32_t bit[kS], d[kS], cls[kS], ef = 0;
    // branch-free per prime (every lane computes every step; the flags
    // decide): a wave almost always has a live lane, so branches only added
    // mask bookkeeping
#pragma unroll
    for (uint32_t u = 0; u < kS; ++u) {
        const uint32_t pu = p[u], pc = pu | (pu == 0);   // pc: a divisor for the idle lanes
        // m = floor(qb0 / p) as in bm_setup: within one, one step each way
        const double pd = (double)pc;
        double inv = (double)__builtin_amdgcn_rcpf((float)pc);
        inv = __builtin_fma(inv, __builtin_fma(-pd, inv, 1.0), inv);
        const double fd = __builtin_floor(qb0d * inv);
        const uint32_t fhi = (uint32_t)(fd * (1.0 / 4294967296.0));
        uint64_t m = ((uint64_t)fhi << 32) | (uint32_t)(fd - (double)fhi * 4294967296.0);
        int64_t r = (int64_t)(h.qb0 - m * pc);
        const bool lo = r < 0, hi = r >= (int64_t)pc;
        m = lo ? m - 1 : (hi ? m + 1 : m);
        r = lo ? r + pc : (hi ? r - (int64_t)pc : r);
        const bool bad = r < 0 || r >= (int64_t)pc;
        const uint32_t r0 = r ? pc - (uint32_t)r : 0u;   // m0 p = qb0 + r0
        const uint64_t m0 = m + (r != 0);
        const uint32_t r30 = ((uint32_t)(m0 >> 32) * 16u + (uint32_t)m0 % 30u) % 30u;
        // a multiple in the span, m0 >= p (smaller ones belong to smaller
        // primes) and a unit mod 30 (else not in the bitmap)
        const bool cand = pu && !bad && r0 <= span && m0 >= pc && ((kUnitMask >> r30) & 1u);
        const bool inb = m0 >= h.m_base && m0 - h.m_base < h.m_span;
        ef |= pu && bad ? 16u : 0u;                      // the quotient more than one off: fail closed
        ef |= cand && !inb ? 8u : 0u;                    // outside the bitmap: fail closed
        const uint64_t y0 = ((m0 - r30 - h.m_base) >> 1) * 0xEEEEEEEEEEEEEEEFull;   // byte of m0
        const uint32_t i = units_below(r30);
        w[u] = y0 >> 3;
        bit[u] = cand && inb ? 8 * (uint32_t)(y0 & 7) + i : 64u;   // 64: no multiple to test
        d[u] = r0 >> 1;
        cls[u] = (cmt[units_below(pc % 30)] >> (3 * i)) & 7;
    }
    if (ef) atomicOr(err, ef);
    uint64_t v[kS];
#pragma unroll
    for (uint32_t u = 0; u < kS; ++u) v[u] = bit[u] < 64 ? __builtin_nontemporal_load(h.mbits + w[u]) : 0ull;
#pragma unroll
    for (uint32_t u = 0; u < kS; ++u) {
        const bool any = bit[u] < 64 && ((v[u] >> bit[u]) & 1ull);
        if (any) file_hit<kMode>(h, df, dl, dq, d[u], cls[u], cnt, hits, cap, err);
        if (kMode == kCount) {
            const uint64_t bal = __ballot(any);
            if ((tid & 63) == 0) any_hit[(k + u) * 4 + (tid >> 6)] = bal;
        }
    }
}

// Bitmap primes with few multipliers in the run's q span (direct form): a
// lean exact scan of every prime finds its multiples q = p m in the span
// with q (so m) a unit mod 30; these candidates are compacted into a
// wave-private LDS queue and resolved together (multiplier, bitmap word and
// bit, class, filing) with every lane busy. Most primes have no candidate
// (about one in eight has one), so the bitmap walk's address work is spent
// on those only. Each wave runs its own rounds: it resolves the previous
// round's queue while the loads of that round's bitmap words and of the next
// round's primes are in flight; no work-group barrier is involved.
//
// The quotient: m_est = round(qb0 / p - 1) from one fused multiply-add
// into the low mantissa bits of 2^52 + m_est (2^52 - 1 is exact, and the sum
// is rounded once, to an integer). With 1 / p from the float reciprocal and
// one Newton step in double (relative error below 2^-44) the computed
// qb0 / p is within Q 2^-44 + 2^11 / p of Q = qb0 / p; for Q < 2^41 and
// p >= 2^20 that is below 0.13, so m_est is floor(Q) or one less and
// r = qb0 - m_est p lies in [0, 2p): one conditional step makes it exact (a
// remainder still out of range sets err bit 4, the run fails). Below
// p = 2^31 the remainder fits 32 bits (kWide: above, 64-bit arithmetic).
constexpr uint32_t kFewPminLog2 = 20;  // primes of the compacted path are >= 2^20
constexpr uint32_t kFewQLog2 = 41;     // and their quotients qb1 / p below 2^41
constexpr uint32_t kFewWave = 256;     // queue slots per wave and round

struct FewCand {
    uint32_t p, d;                     // the prime; half-offset (q - qb0) / 2 of q = p m
    uint32_t m_lo, m_hi_adj;           // m_est: low word, high bits | (m - m_est) << 20
};

// A work-group's queues, one per wave: a round first reads the previous
// round's entries into registers, then scans its primes into the same slots.
struct FewLds {
    FewCand q[4][kFewWave];
};

// x mod 30 for a 32-bit x: a fused double quotient (within 2^-26 of x / 30,
// whose fractional parts are multiples of 1/30) floored after adding 2^-7,
// then the remainder exactly in double (no 32-bit integer multiply or
// division)
__device__ __forceinline__ uint32_t mod30_u32(uint32_t x)
{
    const double xd = (double)x;
    const double fq = __builtin_floor(__builtin_fma(xd, 1.0 / 30.0, 0x1p-7));
    return (uint32_t)__builtin_fma(-30.0, fq, xd);
}

// The per-run constants of the scan.
struct FewRun {
    double qb0d;
    uint32_t qb0_lo, qb0_30;           // low word and residue mod 30 of qb0
    uint64_t span;                     // qb1 - qb0
    uint32_t span32;                   // min(span, 2^32 - 1)
};

// Scans one prime pc (a prime of the packed stream, at least the chunk's
// first; on = false: an idle lane, whose prime lies beyond the run's):
// queues the multiples in the span whose q is a unit mod 30, at
// queue[*n ...]; *n is wave-uniform.
template <uint32_t kM, bool kWide>
__device__ __forceinline__ void few_scan(const HlCtx& h, const FewRun& f, uint32_t pc, bool on, FewCand* queue,
                                         uint32_t& n, bool& bad)
{
    const double pd = (double)pc;
    double inv = (double)__builtin_amdgcn_rcpf((float)pc);
    inv = __builtin_fma(inv, __builtin_fma(-pd, inv, 1.0), inv);
    const uint64_t tb = (uint64_t)__double_as_longlong(__builtin_fma(f.qb0d, inv, 0x1p52 - 1.0));
    uint32_t r;                        // qb0 mod p
    bool step;                         // m_est was floor - 1
    if constexpr (kWide) {
        const uint64_t m_est = tb & ((1ull << 52) - 1);
        uint64_t r64 = h.qb0 - m_est * pc;
        step = r64 >= pc;
        r64 -= step ? pc : 0u;
        bad = bad || (on && r64 >= pc);   // an idle lane's quotient is not used
        r = (uint32_t)r64;
    } else {
        r = f.qb0_lo - (uint32_t)tb * pc;   // exact mod 2^32: the true value is below 2p < 2^32
        step = r >= pc;
        r -= step ? pc : 0u;
        bad = bad || (on && r >= pc);
    }
    const uint32_t r0 = r ? pc - r : 0u;   // the first multiple at or above qb0 is qb0 + r0
    const uint32_t adj = (uint32_t)step + (r != 0);
    uint32_t q30 = mod30_u32(r0) + f.qb0_30;   // q mod 30 of that multiple
    q30 -= q30 >= 30 ? 30u : 0u;
    const uint32_t p30 = kM > 1 ? mod30_u32(pc) : 0u;
#pragma unroll
    for (uint32_t k = 0; k < kM; ++k) {
        const uint64_t off = (uint64_t)r0 + (uint64_t)k * pc;
        const bool in_span = kM == 1 ? r0 <= f.span32 : off <= f.span;
        const bool cand = on && in_span && ((kUnitMask >> q30) & 1u);
        const uint64_t bal = __ballot(cand);
        if (bal) {                             // wave-uniform
            if (cand) {
                const uint32_t at = n + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                queue[at] = FewCand{pc, (uint32_t)(off >> 1), (uint32_t)tb,
                                    ((uint32_t)(tb >> 32) & 0xFFFFFu) | ((adj + k) << 20)};
            }
            n += (uint32_t)__popcll(bal);
        }
        if (kM > 1) {
            q30 += p30;
            q30 -= q30 >= 30 ? 30u : 0u;
        }
    }
}

// kM: multipliers per prime at most (1: primes above the span). A round
// scans kR primes per lane: at most 64 kR kM' candidates per wave (kM' =
// 1 unit among 1 consecutive multiplier, 2 among 4), the wave's queue.
template <uint32_t kM, bool kWide>
__device__ __forceinline__ void few_chunk(const HlCtx& h, const FewRun& f, const gbh_pstream& ps,
                                          uint32_t j0, uint32_t je, uint32_t tid, const uint32_t* df,
                                          const uint32_t* dl, const uint32_t* dq, const uint32_t* cmt, uint32_t* cnt,
                                          FewLds& L, uint32_t* __restrict__ err)
{
    constexpr uint32_t kR = kM == 1 ? 4 : 2;
    constexpr uint32_t kD = kFewWave / 64;    // queued candidates a lane resolves per round, at most
    static_assert(64 * kR * (kM == 1 ? 1 : 2) <= kFewWave, "a round must fit the wave's queue");
    const uint32_t lane = tid & 63, wv = tid >> 6;
    const uint32_t rounds = (je - j0 + 256 * kR - 1) / (256 * kR);
    bool bad = false, outside = false;
    uint32_t p[kR];                    // this round's primes (lanes past je: idle, their primes unused)
    const uint32_t jw = j0 + (tid & ~63u);
#pragma unroll
    for (uint32_t u = 0; u < kR; u += 2) ps_decode2(ps, jw + u * 256, jw + (u + 1) * 256, lane, p[u], p[u + 1]);
    uint32_t prev_n = 0;
    // the last pass only resolves the last round's queue
    for (uint32_t r = 0; r <= rounds; ++r) {
        const FewCand* pq = L.q[wv];
        FewCand c[kD];
        uint64_t w[kD];
        bool live[kD];
#pragma unroll
        for (uint32_t x = 0; x < kD; ++x) {
            live[x] = false;
            w[x] = 0;
            c[x] = FewCand{1u, 0u, 0u, 0u};
            if (64 * x >= prev_n) continue;  // the whole wave's slot x is empty (uniform): nothing to prepare
            const uint32_t e = lane + 64 * x;
            live[x] = e < prev_n;
            c[x] = pq[e < prev_n ? e : 0];
            const uint64_t m = ((((uint64_t)c[x].m_hi_adj & 0xFFFFFu) << 32) | c[x].m_lo) + (c[x].m_hi_adj >> 20);
            // multipliers below p belong to smaller primes
            live[x] = live[x] && m >= c[x].p;
            const bool inb = m >= h.m_base && m - h.m_base < h.m_span;
            outside = outside || (live[x] && !inb);
            live[x] = live[x] && inb;
            // byte y = floor((m - m_base) / 30) of m (m_base is a multiple of
            // 30, so the remainder is m mod 30): m - m_base < 2^42 is exact
            // in double and its fused quotient by 30 within 2^-15 of the true
            // one, whose fractional parts are multiples of 1/30: the floor of
            // the quotient plus 2^-7 is exact
            const uint64_t dm = live[x] ? m - h.m_base : 0ull;
            const double dq = __builtin_fma((double)(uint32_t)(dm >> 32), 4294967296.0, (double)(uint32_t)dm);
            const double fq = __builtin_floor(__builtin_fma(dq, 1.0 / 30.0, 0x1p-7));
            const uint32_t fq_hi = (uint32_t)(fq * (1.0 / 4294967296.0));
            const uint64_t y = ((uint64_t)fq_hi << 32) | (uint32_t)(fq - (double)fq_hi * 4294967296.0);
            const uint32_t r30 = (uint32_t)dm - 30u * (uint32_t)y;
            // one load per slot whether or not it is live (word 0 then): the
            // loads stay unconditional, so waiting for them does not wait for
            // later ones
            w[x] = __builtin_nontemporal_load(h.mbits + (live[x] ? (y >> 3) : 0ull));
            c[x].m_lo = 8 * ((uint32_t)y & 7u) + units_below(r30);   // the bit of m in its word
            c[x].m_hi_adj = r30;
        }
        uint32_t n = 0;
        if (r < rounds) {
            FewCand* q = L.q[wv];
            // the next round's primes: loads in flight during this round's scan
            const uint32_t jn = jw + (r + 1) * 256 * kR;
            PsPair nx[kR / 2];
#pragma unroll
            for (uint32_t u = 0; u < kR; u += 2) nx[u / 2] = ps_load2(ps, jn + u * 256, jn + (u + 1) * 256, lane);
            const uint32_t jr = j0 + r * 256 * kR + tid;
#pragma unroll
            for (uint32_t u = 0; u < kR; ++u) few_scan<kM, kWide>(h, f, p[u], jr + u * 256 < je, q, n, bad);
#pragma unroll
            for (uint32_t u = 0; u < kR; u += 2) ps_finish2(nx[u / 2], p[u], p[u + 1]);
        }
#pragma unroll
        for (uint32_t x = 0; x < kD; ++x) {
            if (live[x] && ((w[x] >> c[x].m_lo) & 1ull)) {
                const uint32_t i = units_below(c[x].m_hi_adj);
                file_hit<kDirect>(h, df, dl, dq, c[x].d, (cmt[units_below(c[x].p % 30)] >> (3 * i)) & 7, cnt, nullptr,
                                  0, err);
            }
        }
        // this wave's queue writes and reads of the round stay in order
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        prev_n = n;
    }
    const uint32_t ef = (bad ? 16u : 0u) | (outside ? 8u : 0u);
    if (ef) atomicOr(err, ef);
}

// One work-group per chunk of primes; two passes over the chunk: count per
// window list, reserve each list's share with one atomic, then write. With
// `stage` slots, the counting pass also keeps the entries in LDS; when they
// all fit, the work-group writes them from there in one burst (its runs of
// each list complete their lines in L2) and skips the second walk.
// Batched ranges: the chunk is walked once per run (a range's consecutive
// segments in the batch), each run with its own q range, lists and prime
// bound; nullptr runs is one run of the whole batch.
// A bitmap prime with p^3 <= qb1 breaks the two-factor rule: fail closed.
// kDirectForm: the direct form (rdesc, region and fills given) or the
// counted / staged forms; kFew: the direct form with the compacted path of
// bitmap primes with few multipliers. One instance each, so none carries
// another's registers (the compacted path's would crowd out the kernels
// that run beside the harvest).
template <bool kDirectForm, bool kFew>
__global__ void __launch_bounds__(256) k_harvest(gbh_batch b, const gbh_pstream ps,
                                                 const uint32_t* __restrict__ chunks,
                                                 const gbh_hl_run* __restrict__ runs, uint32_t nruns, uint32_t nbmax,
                                                 uint32_t t_hl, uint32_t t_hi, uint32_t t_hi_end,
                                                 const uint64_t* __restrict__ mbits,
                                                 uint64_t m_base,
                                                 uint64_t m_span, uint32_t cap, uint32_t stage,
                                                 uint32_t* __restrict__ hits, uint32_t* __restrict__ hit_n,
                                                 const uint32_t* __restrict__ rdesc, uint32_t* __restrict__ region,
                                                 uint32_t* __restrict__ fills, uint32_t* __restrict__ err,
                                                 uint32_t nchunks, uint32_t* __restrict__ next, uint32_t q_off,
                                                 uint32_t hit_queue)
{
    __shared__ uint32_t df[GBH_MAX_BATCH], dl[GBH_MAX_BATCH], dq[GBH_MAX_BATCH * 8], cmt[8], sn, chunk_at;
    __shared__ uint64_t any_hit[GBH_HL_CHUNK / 64 + 4 * (kG > kS ? kG : kS)];   // bitmap primes with a set bit in range (pass 2 walks those)
    __shared__ WheelLds wl;
    extern __shared__ uint32_t cnt[];     // [nseg][8][nwin] counts of a run; with a stage: run counters, the stage
    uint32_t* runc = cnt + nbmax;
    uint32_t* stv = runc + nbmax;
    uint16_t* stb = reinterpret_cast<uint16_t*>(stv + stage);
    const uint32_t tid = threadIdx.x;
    if (tid < 210) { wl.next[tid] = c_w210_next[tid]; wl.idx[tid] = c_w210_idx[tid]; }
    if (tid < 48) wl.inc[tid] = c_w210_inc[tid];
    if (tid < 30) wl.cls[tid] = c_cls_of[tid];
    if (tid < 8) {
        uint32_t v = 0;
        for (uint32_t i = 0; i < 8; ++i) v |= (uint32_t)c_cls_of[(res_of(tid) * res_of(i)) % 30] << (3 * i);
        cmt[tid] = v;
    }
    // persistent form (next given): as many work-groups as the GPU holds at
    // once, each taking chunk after chunk from the device counter, so no
    // chunk waits for a dispatch slot behind a longer one; else chunk =
    // work-group
    for (uint32_t pass = 0;; ++pass) {
    uint32_t kc;
    if (next) {
        __syncthreads();                      // the previous chunk is done with chunk_at and the LDS tables
        if (tid == 0) chunk_at = atomicAdd(next, 1u);
        __syncthreads();
        kc = chunk_at;
        if (kc >= nchunks) break;
    } else {
        if (pass) break;
        kc = blockIdx.x;
    }
    const uint32_t j0 = chunks[kc], j1 = chunks[kc + 1];
    // beyond the pass-2 flags, or not at a group of the packed stream: fail closed
    if (j1 - j0 > GBH_HL_CHUNK || (j0 & (GBH_PS_GROUP - 1)) || j1 > ps.n) {
        if (tid == 0) atomicOr(err, 1u);
        continue;
    }
    // the chunk's first prime (a group's base) and a bound on its last (the
    // base of the group after the last prime's)
    const uint32_t pfirst = ps.base[j0 / GBH_PS_GROUP];
    // a chunk is all wheel primes or all bitmap primes (the host splits at t_hl, t_hi and t_hi_end)
    const bool wheel = j0 < j1 && (pfirst < t_hl || (pfirst >= t_hi && pfirst < t_hi_end));
    for (uint32_t r = 0; r < nruns; ++r) {
        // every condition below is uniform over the work-group
        const uint32_t seg0 = runs ? runs[r].seg0 : 0u, rn = runs ? runs[r].nseg : b.nseg;
        const uint32_t je = runs ? min(j1, runs[r].nprimes) : j1;
        constexpr bool direct = kDirectForm;
        const uint32_t nb = direct ? rn * 8 : rn * 8 * b.nwin;   // the run's counters
        if (rn == 0 || seg0 + rn > b.nseg || nb > nbmax) {       // a run beyond the contract: fail closed
            if (tid == 0) atomicOr(err, 1u);
            continue;
        }
        if (j0 >= je) {                       // no prime of the chunk in this run: empty regions
            if (direct) {
                uint32_t* fl = fills + (size_t)kc * (b.nseg * 8) + seg0 * 8;
                for (uint32_t i = tid; i < nb; i += 256) fl[i] = 0;
            }
            continue;
        }
        __syncthreads();                      // the previous run is done with the LDS tables
        const gbh_geom* gs = b.geoms + seg0;
        const uint64_t qb0 = gs[0].q_first;
        for (uint32_t s = tid; s < rn; s += 256) {
            df[s] = (uint32_t)((gs[s].q_first - qb0) >> 1);
            dl[s] = (uint32_t)((gs[s].q_last - qb0) >> 1);
        }
        for (uint32_t i = tid; i < rn * 8; i += 256) dq[i] = (uint32_t)((gs[i / 8].Q[i % 8] - qb0) >> 1);
        for (uint32_t i = tid; i < nb; i += 256) {
            cnt[i] = 0;
            if (stage) runc[i] = 0;
        }
        if (tid == 0) sn = 0;
        __syncthreads();
        HlCtx h;
        h.stv = stage ? stv : nullptr;
        h.stb = stb;
        h.sn = &sn;
        h.scap = stage;
        h.region = direct ? region + rdesc[2 * kc] + (size_t)seg0 * 8 * rdesc[2 * kc + 1] : nullptr;
        h.rcap = direct ? rdesc[2 * kc + 1] : 0u;
        h.qb0 = qb0;
        h.dmax = dl[rn - 1];
        h.hs_magic = (rn > 1 && df[1] > df[0]) ? (uint32_t)((1ull << 32) / (df[1] - df[0] + 1)) : 0u;
        h.nseg = rn;
        h.nwin = b.nwin;
        h.lw = 31 - __clz(b.wbits);
        h.wmask = b.wbits - 1;
        h.t_hl = t_hl;
        h.mbits = mbits;
        h.m_base = m_base;
        h.m_span = m_span;
        const uint64_t qb1 = qb0 + 2ull * h.dmax;
        const double qb0d = (double)qb0;
        if (!wheel) {                         // the chunk's smallest prime carries the two-factor rule
            const uint64_t p = pfirst, p2 = p * p;
            if (__umul64hi(p2, p) == 0 && p2 * p <= qb1) {
                if (tid == 0) atomicOr(err, 4u);
                continue;
            }
        }
        // every prime of the chunk above the run's q span: the lean form
        const bool single = !wheel && (uint64_t)pfirst > qb1 - qb0;
        uint32_t* rh = hits + (size_t)seg0 * 8 * b.nwin * cap;   // the run's lists
        uint32_t* rhn = hit_n + (size_t)seg0 * 8 * b.nwin;
        if constexpr (kDirectForm) {   // one walk into the work-group's regions, then their fill counts
            // bitmap primes with at most 4 multipliers in the span, inside the
            // compacted path's quotient bounds (uniform over the chunk)
            const uint64_t pmin = pfirst;
            const bool few = kFew && q_off && !wheel && pmin >= (1u << kFewPminLog2) && (qb1 >> kFewQLog2) < pmin &&
                             qb1 - qb0 < 4 * pmin;
            const bool wide = ps.base[(je - 1) / GBH_PS_GROUP + 1] >= (1u << 31);   // (an upper bound)
            if (wheel) {
                for (uint32_t jb = j0; jb < je; jb += 256) {
                    const uint32_t p = ps_decode1(ps, jb + (tid & ~63u), tid & 63);
                    if (jb + tid < je) walk_wheel<kDirect>(h, qb0d, p, df, dl, dq, wl, cnt, rh, cap, err);
                }
            } else if (kFew && few) {
                // the queues: the dynamic LDS from byte q_off
                FewLds& few_lds = *reinterpret_cast<FewLds*>(reinterpret_cast<char*>(cnt) + q_off);
                const uint64_t span = qb1 - qb0;
                const FewRun f{qb0d, (uint32_t)qb0, (uint32_t)(qb0 % 30), span,
                               (uint32_t)min(span, (uint64_t)0xFFFFFFFFu)};
                if (single && wide) few_chunk<1, true>(h, f, ps, j0, je, tid, df, dl, dq, cmt, cnt, few_lds, err);
                else if (single) few_chunk<1, false>(h, f, ps, j0, je, tid, df, dl, dq, cmt, cnt, few_lds, err);
                else if (wide) few_chunk<4, true>(h, f, ps, j0, je, tid, df, dl, dq, cmt, cnt, few_lds, err);
                else few_chunk<4, false>(h, f, ps, j0, je, tid, df, dl, dq, cmt, cnt, few_lds, err);
            } else if (kFew && !single && q_off && hit_queue) {
                // (the instance with the compacted path only: its registers
                // are spent already, the other instance keeps its waves)
                HitQLds& hq = *reinterpret_cast<HitQLds*>(reinterpret_cast<char*>(cnt) + q_off);
                uint32_t n = 0;               // this wave's queued hits
                for (uint32_t k = 0; j0 + k * 256 < je; k += kGQ)
                    bm_group_q(h, qb0d, ps, j0, je, k, tid, df, dl, dq, cmt, cnt, hq, n, err);
                if (n) hitq_flush(h, df, dl, dq, cnt, err, hq.d[tid >> 6], hq.c[tid >> 6], n, tid & 63);
            } else {
                for (uint32_t k = 0; j0 + k * 256 < je; k += single ? kS : kG) {
                    if (single) bm_single<kDirect>(h, qb0d, ps, j0, je, k, tid, df, dl, dq, cmt, cnt, any_hit, rh, cap, err);
                    else bm_group<kDirect>(h, qb0d, ps, j0, je, k, tid, df, dl, dq, cmt, cnt, any_hit, rh, cap, err);
                }
            }
            __syncthreads();
            uint32_t* fl = fills + (size_t)kc * (b.nseg * 8) + seg0 * 8;
            for (uint32_t i = tid; i < nb; i += 256) fl[i] = min(cnt[i], h.rcap);
            continue;
        }
        if (wheel) {
            for (uint32_t jb = j0; jb < je; jb += 256) {
                const uint32_t p = ps_decode1(ps, jb + (tid & ~63u), tid & 63);
                if (jb + tid < je) walk_wheel<kCount>(h, qb0d, p, df, dl, dq, wl, cnt, rh, cap, err);
            }
        } else {
            for (uint32_t k = 0; j0 + k * 256 < je; k += single ? kS : kG) {
                if (single) bm_single<kCount>(h, qb0d, ps, j0, je, k, tid, df, dl, dq, cmt, cnt, any_hit, rh, cap, err);
                else bm_group<kCount>(h, qb0d, ps, j0, je, k, tid, df, dl, dq, cmt, cnt, any_hit, rh, cap, err);
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < nb; i += 256) {
            const uint32_t c = cnt[i];
            cnt[i] = c ? atomicAdd(&rhn[i], c) : 0u;
        }
        __syncthreads();
        if (stage && sn <= stage) {          // every entry is in the stage: write it from there
            const uint32_t n = sn;
            for (uint32_t i = tid; i < n; i += 256) {
                const uint32_t bk = stb[i];
                const uint32_t pos = cnt[bk] + atomicAdd(&runc[bk], 1u);
                if (pos < cap) rh[(size_t)bk * cap + pos] = stv[i];
                else atomicOr(err, 1u);
            }
            continue;
        }
        if (wheel) {
            for (uint32_t jb = j0; jb < je; jb += 256) {
                const uint32_t p = ps_decode1(ps, jb + (tid & ~63u), tid & 63);
                if (jb + tid < je) walk_wheel<kWrite>(h, qb0d, p, df, dl, dq, wl, cnt, rh, cap, err);
            }
        } else {
            for (uint32_t k = 0; j0 + k * 256 < je; k += single ? kS : kG) {
                if (single) bm_single<kWrite>(h, qb0d, ps, j0, je, k, tid, df, dl, dq, cmt, cnt, any_hit, rh, cap, err);
                else bm_group<kWrite>(h, qb0d, ps, j0, je, k, tid, df, dl, dq, cmt, cnt, any_hit, rh, cap, err);
            }
        }
    }
    }
}

// Direct harvest, second step: the (segment, class) regions every chunk's
// work-group filled are split into the window lists. Work-group (bk, g):
// bucket bk = s * 8 + c of the batch, the chunks [groups[g], groups[g + 1])
// (at most GBH_PART_CHUNKS, their regions at most kCap entries in all:
// planned by the host). The entries are read once into LDS (the chunks'
// fill counts scanned so every thread's loads are independent) and counted
// per window; one reservation per window list; then the writes from LDS
// (each window list's share lands in consecutive slots).
template <uint32_t kCap>
__global__ void __launch_bounds__(512) k_partition(const uint32_t* __restrict__ rdesc,
                                                   const uint32_t* __restrict__ region,
                                                   const uint32_t* __restrict__ fills,
                                                   const uint32_t* __restrict__ groups, uint32_t nbb, uint32_t nwin,
                                                   uint32_t wbits, uint32_t cap, uint32_t* __restrict__ hits,
                                                   uint32_t* __restrict__ hit_n, uint32_t* __restrict__ err)
{
    __shared__ uint32_t ent[kCap];
    __shared__ uint32_t wc[256], wbase[256], pre[GBH_PART_CHUNKS + 1], rb[GBH_PART_CHUNKS];
    const uint32_t bk = blockIdx.x, tid = threadIdx.x;
    const uint32_t k0 = groups[blockIdx.y], k1 = groups[blockIdx.y + 1];
    // a group that is empty, reversed or longer than the LDS tables is
    // beyond the host's plan: its regions would go unread, so fail closed
    if (k1 <= k0 || k1 - k0 > GBH_PART_CHUNKS) {
        if (tid == 0) atomicOr(err, 1u);
        return;
    }
    const uint32_t nk = k1 - k0;
    const uint32_t lw = 31 - __clz(wbits);
    for (uint32_t i = tid; i < nwin; i += 512) wc[i] = 0;
    if (tid < 64) {   // one wave: fill counts, region bases and their exclusive scan
        uint32_t n = 0;
        if (tid < nk) {
            const uint32_t k = k0 + tid;
            n = min(fills[(size_t)k * nbb + bk], rdesc[2 * k + 1]);   // within the region whatever happened
            rb[tid] = rdesc[2 * k] + bk * rdesc[2 * k + 1];
        }
        uint32_t tot = 0;
        const uint32_t x = wave_excl_scan(n, &tot);
        pre[tid] = x;
        if (tid == 0) pre[nk] = tot;
    }
    __syncthreads();
    const uint32_t total = pre[nk];
    if (total > kCap) {                      // beyond the host's plan: fail closed
        if (tid == 0) atomicOr(err, 1u);
        return;
    }
    // the entries, kPartLoads per thread in flight at a time
    constexpr uint32_t kPartLoads = 8;
    for (uint32_t f0 = 0; f0 < total; f0 += 512 * kPartLoads) {
        uint32_t v[kPartLoads];
#pragma unroll
        for (uint32_t u = 0; u < kPartLoads; ++u) {
            const uint32_t f = f0 + u * 512 + tid;
            v[u] = 0xFFFFFFFFu;
            if (f < total) {
                uint32_t lo = 0, hi = nk;    // the chunk holding f: the last k with pre[k] <= f
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (pre[mid] <= f) lo = mid;
                    else hi = mid;
                }
                v[u] = __builtin_nontemporal_load(region + rb[lo] + (f - pre[lo]));   // read once
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kPartLoads; ++u) {
            const uint32_t f = f0 + u * 512 + tid;
            if (f >= total) continue;
            uint32_t t = v[u];
            if ((t >> lw) >= nwin) {
                atomicOr(err, 1u);
                t = 0xFFFFFFFFu;
            } else {
                atomicAdd(&wc[t >> lw], 1u);
            }
            ent[f] = t;
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < nwin; i += 512) {   // one reservation per window list
        const uint32_t c = wc[i];
        wbase[i] = c ? atomicAdd(&hit_n[(size_t)bk * nwin + i], c) : 0u;
        wc[i] = 0;
    }
    __syncthreads();
    for (uint32_t f = tid; f < total; f += 512) {
        const uint32_t t = ent[f];
        if (t == 0xFFFFFFFFu) continue;
        const uint32_t w = t >> lw, pos = wbase[w] + atomicAdd(&wc[w], 1u);
        if (pos < cap) hits[((size_t)bk * nwin + w) * cap + pos] = t & (wbits - 1);
        else atomicOr(err, 1u);
    }
}

// 8 x 8 bit transpose: bit 8 r + c <-> bit 8 c + r (three delta swaps)
__device__ __forceinline__ uint64_t transpose8(uint64_t x)
{
    uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
    x ^= t ^ (t << 7);
    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
    x ^= t ^ (t << 14);
    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
    x ^= t ^ (t << 28);
    return x;
}

// Multiplier bitmap from a batch of class maps over m: segment s holds the
// bytes y in [s nc, (s + 1) nc) of the bitmap (all eight class arrays start
// at the same byte), byte y holds the eight class bits of y.
__global__ void __launch_bounds__(256) k_mtrans(gbh_batch b, uint32_t nc, uint64_t* __restrict__ mbits,
                                                uint64_t word0, uint64_t mwords)
{
    const uint32_t x = blockIdx.x * 256 + threadIdx.x;   // class word of the segment
    const uint32_t s = blockIdx.y;
    if (x >= nc / 64) return;
    const uint64_t* map = seg_map(b, s);
    uint64_t cwd[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) cwd[c] = map[(size_t)c * b.cw + x];
    const uint64_t ow = word0 + (uint64_t)s * (nc / 8) + 8ull * x;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        uint64_t v = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) v |= ((cwd[i] >> (8 * j)) & 0xFFull) << (8 * i);
        if (ow + j < mwords) mbits[ow + j] = transpose8(v);
    }
}

// ---------------------------------------------------------------- stage 3
// Offset from wb0 of the first mark >= wb0 of a prime whose first class bit
// is t (t may be below wb0): (t - wb0) mod p. Exact integer division, or
// (first_off_big) a float quotient (within a few units of d / p for d <
// 2^31, p >= 67) corrected by the remainder.
__device__ __forceinline__ uint32_t first_off(uint32_t t, uint32_t wb0, uint32_t p)
{
    if (t >= wb0) return t - wb0;
    const uint32_t d = wb0 - t;
    const uint32_t r = d % p;
    return r ? p - r : 0u;
}

__device__ __forceinline__ uint32_t first_off_big(uint32_t t, uint32_t wb0, uint32_t p, float rp)
{
    if (t >= wb0) return t - wb0;
    const uint32_t d = wb0 - t;
    // d < 2^31, 67 <= p < 2^21: the float quotient is within 7 of d / p (within
    // one and a bit for p >= 2^16), so |r| < 2^24 (exact in float); a second
    // float quotient of r is within one of r / p; then one step each way. No
    // correction loop: such a loop is compiled into an integer division.
    int32_t r = (int32_t)(d - (uint32_t)((float)d * rp) * p);
    r -= (int32_t)((float)r * rp) * (int32_t)p;
    if (r < 0) r += (int32_t)p;
    else if (r >= (int32_t)p) r -= (int32_t)p;
    return r ? p - (uint32_t)r : 0u;
}

// Sets the bits bt, bt + step, ... below wbits of an LDS window (composites
// as set bits), two per iteration.
__device__ __forceinline__ void mark_run(uint32_t* win32, uint32_t bt, uint32_t step, uint32_t wbits)
{
    for (; bt + step < wbits; bt += 2 * step) {
        atomicOr(&win32[bt >> 5], 1u << (bt & 31));
        const uint32_t b2 = bt + step;
        atomicOr(&win32[b2 >> 5], 1u << (b2 & 31));
    }
    if (bt < wbits) atomicOr(&win32[bt >> 5], 1u << (bt & 31));
}

__global__ void __launch_bounds__(256) k_window(gbh_batch b, const uint32_t* __restrict__ primes, uint32_t np,
                                                uint32_t ncoop, const uint32_t* __restrict__ i0tab,
                                                const uint32_t* __restrict__ hits, const uint32_t* __restrict__ hit_n,
                                                uint32_t cap, uint32_t phases)
{
    extern __shared__ uint64_t win[];
    uint32_t* win32 = reinterpret_cast<uint32_t*>(win);
    const uint32_t c = blockIdx.y, s = blockIdx.z;
    const gbh_geom& g = b.geoms[s];
    if (blockIdx.x >= g.nwin) return;
    const uint32_t nwords = b.wbits >> 6;
    const uint64_t gw0 = (uint64_t)blockIdx.x * nwords;     // first word of the window in class c
    const uint32_t wb0 = (uint32_t)(gw0 * 64);              // first bit (class arrays < 2^31 bits)
    const uint32_t tid = threadIdx.x;
    const uint32_t* i0 = i0tab + ((size_t)s * 8 + c) * np;
    const bool sieve = phases & 1, apply = (phases & 2) && hits;
    uint64_t* dst = seg_map(b, s) + (size_t)c * b.cw + gw0;
    // the words of the window that hold bits of the class array (a short
    // segment's window, or a segment's last one, ends early): only they are
    // sieved and stored; the rest of the window is stored as zeros (no
    // primes), which every reader masks off anyway
    const uint32_t ncc = g.nc[c];
    const uint32_t vwords = ncc > wb0 ? min(nwords, (ncc - wb0 + 63) >> 6) : 0u;
    const uint32_t vlim = vwords * 64;                      // marks below it
    const uint32_t wb1 = wb0 + vlim;

    // In LDS a window holds the composites as set bits (marks are ORs; the
    // store inverts it back to the map's primes-as-ones).
    if (!sieve) {   // the hit lists only: the window as the sieve phase stored it
        uint32_t x = tid;
        for (; x + 3 * 256 < vwords; x += 4 * 256) {   // four loads in flight per thread
            uint64_t v[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) v[u] = dst[x + u * 256];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) win[x + u * 256] = ~v[u];
        }
        for (; x < vwords; x += 256) win[x] = ~dst[x];
    } else {        // 1) mask stage: every multiple of 7..61 set by whole-word masks
        uint32_t ph[GBH_NSMALL];
#pragma unroll
        for (int k = 0; k < GBH_NSMALL; ++k) {
            const uint32_t p = kSmall[k];
            const uint32_t a = (uint32_t)(((gw0 + tid) * 64) % p);
            ph[k] = (g.s0[c][k] + p - a) % p;
        }
        for (uint32_t x = tid; x < vwords; x += 256) {
            uint64_t m = 0;
#pragma unroll
            for (int k = 0; k < GBH_NSMALL; ++k) {
                m |= kPat[k] << ph[k];
                const uint32_t p = kSmall[k];
                const uint32_t st = kStep256[k];                 // (64 * 256) mod p
                ph[k] = ph[k] >= st ? ph[k] - st : ph[k] + p - st;
            }
            if (gw0 + x == 0) m &= ~g.restore[c];
            win[x] = m;
        }
    }
    __syncthreads();

    // 2) window primes below wbits / 128 (at least 128 marks each): one
    //    wave per prime, lane l takes marks l, l + 64, ...
    const uint32_t lane = tid & 63, wv = tid >> 6;
    np = sieve ? min(np, g.np_win) : 0u;   // the segment's primes (larger ones have no multiple in it)
    ncoop = min(ncoop, np);
    {   // the wave's primes are uniform: scalar loads, the next one's issued
        // before the current one's marks
        const uint32_t w0 = __builtin_amdgcn_readfirstlane(wv);
        uint32_t t0n = w0 < ncoop ? i0[w0] : 0xFFFFFFFFu, pn = w0 < ncoop ? primes[w0] : 0u;
        for (uint32_t j = w0; j < ncoop; j += 4) {
            const uint32_t t0 = t0n, p = pn;
            if (j + 4 < ncoop) {
                t0n = i0[j + 4];
                pn = primes[j + 4];
            }
            if (t0 == 0xFFFFFFFFu) continue;
            mark_run(win32, first_off(t0, wb0, p) + lane * p, 64 * p, vlim);
        }
    }
    // 3) the other window primes: one prime per lane, kWL of them loaded at
    //    a time
    constexpr uint32_t kWL = 4;
    for (uint32_t j0 = ncoop + tid; j0 < np; j0 += 256 * kWL) {
        uint32_t t0[kWL], pr[kWL];
#pragma unroll
        for (uint32_t u = 0; u < kWL; ++u) {
            const uint32_t j = j0 + 256 * u;
            t0[u] = j < np ? i0[j] : 0xFFFFFFFFu;
            pr[u] = j < np ? primes[j] : 1u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kWL; ++u) {
            if (t0[u] == 0xFFFFFFFFu || t0[u] >= wb1) continue;
            const uint32_t p = pr[u];
            mark_run(win32, first_off_big(t0[u], wb0, p, __builtin_amdgcn_rcpf((float)p)), p, vlim);
        }
    }
    // 4) the large primes' hits of this window
    if (apply) {
        const uint32_t bucket = (s * 8 + c) * b.nwin + blockIdx.x;
        const uint32_t n = min(hit_n[bucket], cap);
        const uint32_t* hl = hits + (size_t)bucket * cap;
        uint32_t i = tid;
        for (; i + 3 * 256 < n; i += 4 * 256) {   // four loads in flight per thread
            uint32_t bt[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) bt[u] = hl[i + u * 256];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) atomicOr(&win32[bt[u] >> 5], 1u << (bt[u] & 31));
        }
        for (; i < n; i += 256) {
            const uint32_t bt = hl[i];
            atomicOr(&win32[bt >> 5], 1u << (bt & 31));
        }
    }
    __syncthreads();

    // 5) coalesced store of the window (primes as set bits again)
    for (uint32_t x = tid; x < nwords; x += 256) dst[x] = x < vwords ? ~win[x] : 0ull;
}

// ---------------------------------------------------------------- checker
struct ClassEvens {
    uint64_t n0;      // first checked even of the class
    uint32_t cnt;     // checked evens of the class
    uint32_t i0;      // even index of n0 in the segment
};

__device__ __forceinline__ ClassEvens class_evens(const gbh_geom& g, uint32_t cr)
{
    ClassEvens e;
    const uint64_t n_gpu = g.n_gpu;
    const uint32_t r = (uint32_t)(n_gpu % 30);
    e.n0 = n_gpu + (2 * cr + 30 - r) % 30;
    e.cnt = e.n0 <= g.seg_end ? (uint32_t)((g.seg_end - e.n0) / 30 + 1) : 0u;
    e.i0 = (uint32_t)((e.n0 - g.seg_start) >> 1);
    return e;
}

// Phase A layout: a work-group stages words [w0, w0 + kRowW) of the eight
// class arrays in LDS as two planes of 32-bit halves, the low dwords of
// every row first, then the high dwords (lane i reads dword x + i of a plane:
// unit stride across the wave, no bank conflicts). Lane i of wave v takes
// class-array word w0 + 63 v + i of every even class (64 evens of the class,
// one per bit); lane 63 only lends its first dword to lane 62. For candidate
// j the 64 bits the lane needs start at bit sh of dword a, then dwords b and
// c follow: a and b are read at plane offsets fa + 63 v + i and fb + 63 v + i
// (a low half then the same word's high half, or a high half then the next
// word's low half), c is lane i + 1's a, and two funnel shifts give the 64
// bits.
constexpr uint32_t kWaveWords = 63;
constexpr uint32_t kBlockWords = kWaveWords * (GBH_CHECK_WG / 64);   // 252 words of every class per work-group
constexpr uint32_t kRowW = kBlockWords + 12;                         // staged words per class row (offsets < 12 words)
constexpr uint32_t kPlane = GBH_CLASSES * kRowW;                     // dwords per plane (low / high halves)
constexpr uint32_t kPbuf = 2048;                                     // pending evens staged per work-group
// a candidate's offset is below (9781 + 58) / 30 bits (n0 <= n_gpu + 28 <= q_first + 9810)
static_assert((9781 + 58) / 30 / 64 + kBlockWords + 2 <= kRowW, "staged rows too short for the offsets");
static_assert(2 * kPlane < (1u << 13), "plane offsets must fit 13 bits");

// Per segment and even class cr, candidate j:
//   coff[(s * 15 + cr) * GBH_NCAND + j] = class << 16 | bit offset (phase B)
//   coffa[(s * 15 + cr) * GBH_NCAND_PAD + j] = fa | fb << 13 | shift << 26 (phase A)
// where the bit offset is (n0 - p - Q[class]) / 30 for the class's first
// checked even n0, and fa / fb the staged plane offsets of the dwords a / b
// that hold it (word off / 64 of the class row, its low or high half).
__global__ void __launch_bounds__(256) k_coff(gbh_batch b, uint32_t* __restrict__ coff, uint32_t* __restrict__ coffa)
{
    const uint32_t cr = blockIdx.y, s = blockIdx.z;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= GBH_NCAND_PAD) return;
    const gbh_geom& g = b.geoms[s];
    if (g.n_gpu > g.seg_end) return;
    uint32_t* pa = coffa + ((size_t)s * 15 + cr) * GBH_NCAND_PAD;
    if (j >= c_cand_n[cr]) {
        pa[j] = 0;
        return;
    }
    const ClassEvens e = class_evens(g, cr);
    const uint32_t p = c_cand_p[cr * GBH_NCAND + j];
    const uint32_t c = c_cand_c[cr * GBH_NCAND + j];
    const uint32_t off = (uint32_t)((e.n0 - p - g.Q[c]) / 30);   // exact and >= 0 (host contract on n_gpu)
    coff[((size_t)s * 15 + cr) * GBH_NCAND + j] = (c << 16) | off;
    const uint32_t x = c * kRowW + (off >> 6);
    const uint32_t fa = (off & 32) ? kPlane + x : x, fb = (off & 32) ? x + 1 : kPlane + x;
    pa[j] = fa | (fb << 13) | ((off & 31) << 26);
}

// Sum over the wave, in lane 63: inclusive row sums by DPP row_shr 1 2 4 8,
// then the row totals folded by row_bcast:15 and row_bcast:31.
__device__ __forceinline__ uint32_t wave_total63(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x111, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x112, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x114, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x118, 0xF, 0xF, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);
    return v;
}

// lane i receives lane i + 1's value (DPP wave_shl:1; lane 63 receives 0)
__device__ __forceinline__ uint32_t from_next_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xF, 0xF, true);
}

template <bool kHist>
__global__ void __launch_bounds__(GBH_CHECK_WG) k_check_a(gbh_batch b, const uint32_t* __restrict__ coffa,
                                                          uint32_t ka, uint32_t* __restrict__ pending, uint32_t pcap,
                                                          gbh_segout* __restrict__ outs)
{
    __shared__ uint32_t st[2 * kPlane];
    __shared__ uint32_t pbuf[kPbuf];
    __shared__ uint32_t hl[kHist ? GBH_EVEN_CLASSES * GBH_KA_MAX : 1];
    __shared__ uint32_t pn, pres, pbase;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, s = blockIdx.y;
    const gbh_geom& g = b.geoms[s];
    if (g.n_gpu > g.seg_end || (g.hist_on != 0) != kHist) return;
    gbh_segout* out = outs + s;
    const uint64_t* map = seg_map(b, s);
    const uint32_t w0 = blockIdx.x * kBlockWords;
    // stage the rows: every load issued before the first store
    {
        constexpr uint32_t kN = GBH_CLASSES * kRowW, kPer = (kN + GBH_CHECK_WG - 1) / GBH_CHECK_WG;
        uint64_t v[kPer];
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t i = tid + k * GBH_CHECK_WG;
            const uint32_t c = i / kRowW, x = i % kRowW;
            v[k] = (i < kN && w0 + x < b.cw) ? map[(size_t)c * b.cw + w0 + x] : 0ull;
        }
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            const uint32_t i = tid + k * GBH_CHECK_WG;
            if (i < kN) {                    // i = c * kRowW + x
                st[i] = (uint32_t)v[k];
                st[kPlane + i] = (uint32_t)(v[k] >> 32);
            }
        }
    }
    if (kHist)
        for (uint32_t i = tid; i < GBH_EVEN_CLASSES * GBH_KA_MAX; i += GBH_CHECK_WG) hl[i] = 0;
    if (tid == 0) { pn = 0; pres = 0; }
    __syncthreads();

    const uint32_t word = w0 + wv * kWaveWords + lane;     // this lane's word of every class array
    const uint32_t* sb = st + wv * kWaveWords + lane;
    uint32_t resolved = 0;
    // the work-group's even classes: all 15, or its share of a split
    const uint32_t ncr = GBH_EVEN_CLASSES / gridDim.z, cr0 = blockIdx.z * ncr;
    for (uint32_t cr = cr0; cr < cr0 + ncr; ++cr) {
        const ClassEvens e = class_evens(g, cr);
        const uint32_t k0 = word * 64;
        uint64_t valid = 0;
        if (lane < kWaveWords && k0 < e.cnt) valid = (e.cnt - k0 >= 64) ? ~0ull : ((1ull << (e.cnt - k0)) - 1);
        const uint32_t v0 = (uint32_t)valid, v1 = (uint32_t)(valid >> 32);
        uint32_t u0 = v0, u1 = v1;                        // unresolved evens of the word
        uint32_t nu = kHist ? (uint32_t)__popc(u0) + (uint32_t)__popc(u1) : 0u;   // how many
        // ka is a multiple of 16 and at most every class's useful candidates
        const uint32_t* ct = coffa + ((size_t)s * 15 + cr) * GBH_NCAND_PAD;
        for (uint32_t j0 = 0; j0 < ka; j0 += 16) {
            if (kHist) {
                // evens first resolved by candidate j0 + u (the drop in the
                // unresolved count), two counts per register (16 bits
                // each: at most 64 x 64 per wave)
                uint32_t hc[8];
#pragma unroll
                for (uint32_t u = 0; u < 16; ++u) {
                    const uint32_t ent = ct[j0 + u], sh = ent >> 26;
                    const uint32_t a = sb[ent & 0x1FFF], bw = sb[(ent >> 13) & 0x1FFF], cw = from_next_lane(a);
                    const uint32_t x0 = __builtin_amdgcn_alignbit(bw, a, sh);
                    const uint32_t x1 = __builtin_amdgcn_alignbit(cw, bw, sh);
                    u0 &= ~x0;
                    u1 &= ~x1;
                    const uint32_t left = (uint32_t)__popc(u0) + (uint32_t)__popc(u1);
                    const uint32_t h = nu - left;
                    nu = left;
                    hc[u / 2] = (u & 1) ? (hc[u / 2] | (h << 16)) : h;
                }
                // the sixteen wave totals by a halving exchange: each step
                // a lane keeps half of its counts, adds its partner's copy
                // of that half and hands over the other half. Partners:
                // lane ^ 32 and ^ 16, then inside a row of 16 lane ^ 15
                // (DPP row_mirror) and ^ 7 (row_half_mirror), each
                // differing in the bit that picks the half. The first two
                // steps are register swaps across rows: permlane32_swap(x,
                // y) trades x's upper half-wave for y's lower one, so the
                // sum of its two results is x summed over lanes l and
                // l + 32 in the lower half and y so summed in the upper;
                // permlane16_swap does the same for odd and even rows.
                // After four steps lane l holds count (l >> 2) & 15 summed
                // over eight lanes of its row; quad_perm ^ 1 and ^ 2 add
                // the rest. Fields stay below 2^10 before the unpacking (64
                // per lane, doubling each step).
                const bool b3 = lane & 8, b2 = lane & 4;
                uint32_t h4[4], h2[2];
#pragma unroll
                for (uint32_t k = 0; k < 4; ++k) {
                    const auto sw = __builtin_amdgcn_permlane32_swap(hc[k], hc[k + 4], false, false);
                    h4[k] = sw[0] + sw[1];
                }
#pragma unroll
                for (uint32_t k = 0; k < 2; ++k) {
                    const auto sw = __builtin_amdgcn_permlane16_swap(h4[k], h4[k + 2], false, false);
                    h2[k] = sw[0] + sw[1];
                }
                const uint32_t h1 = (b3 ? h2[1] : h2[0]) +
                                    (uint32_t)__builtin_amdgcn_mov_dpp((int)(b3 ? h2[0] : h2[1]), 0x140, 0xF, 0xF, false);
                uint32_t tv = (b2 ? h1 >> 16 : h1 & 0xFFFFu) +
                              (uint32_t)__builtin_amdgcn_mov_dpp((int)(b2 ? h1 & 0xFFFFu : h1 >> 16), 0x141, 0xF, 0xF, false);
                tv += (uint32_t)__builtin_amdgcn_mov_dpp((int)tv, 0xB1, 0xF, 0xF, false);
                tv += (uint32_t)__builtin_amdgcn_mov_dpp((int)tv, 0x4E, 0xF, 0xF, false);
                if ((lane & 3) == 0 && tv) atomicAdd(&hl[cr * GBH_KA_MAX + j0 + (lane >> 2)], tv);
            } else {
                uint32_t r0 = 0, r1 = 0;
#pragma unroll
                for (uint32_t u = 0; u < 16; ++u) {
                    const uint32_t ent = ct[j0 + u], sh = ent >> 26;
                    const uint32_t a = sb[ent & 0x1FFF], bw = sb[(ent >> 13) & 0x1FFF], cw = from_next_lane(a);
                    r0 |= __builtin_amdgcn_alignbit(bw, a, sh);
                    r1 |= __builtin_amdgcn_alignbit(cw, bw, sh);
                }
                u0 &= ~r0;
                u1 &= ~r1;
            }
            if (!__any(u0 | u1)) break;
        }
        const uint64_t unres = ((uint64_t)u1 << 32) | u0;
        resolved += (uint32_t)__popcll(valid & ~unres);
        // pending evens: staged in LDS, one LDS atomic per wave
        const uint32_t cntp = (uint32_t)__popcll(unres);
        uint32_t tot = 0;
        const uint32_t pre = wave_excl_scan(cntp, &tot);
        if (tot) {
            uint32_t base = 0;
            if (lane == 63) base = atomicAdd(&pn, tot);
            base = __shfl(base, 63, 64);
            uint32_t pos = base + pre;
            uint64_t m = unres;
            while (m) {
                const uint32_t bt = (uint32_t)__ffsll((long long)m) - 1;
                m &= m - 1;
                const uint32_t idx = e.i0 + 15u * (k0 + bt);
                if (pos < kPbuf) {
                    pbuf[pos] = idx;
                } else {                                   // overflow: straight to the segment's list
                    const uint32_t q = atomicAdd(&out->pending_n, 1u);
                    if (q < pcap) pending[(size_t)s * pcap + q] = idx;
                }
                ++pos;
            }
        }
    }
    const uint32_t rs = wave_sum(resolved);
    if (lane == 0 && rs) atomicAdd(&pres, rs);
    __syncthreads();
    const uint32_t np = min(pn, kPbuf);
    if (tid == 0) {
        pbase = np ? atomicAdd(&out->pending_n, np) : 0u;
        if (pres) atomicAdd(&out->acct[0], pres);
    }
    __syncthreads();
    uint32_t* pend = pending + (size_t)s * pcap;
    for (uint32_t i = tid; i < np; i += GBH_CHECK_WG)
        if (pbase + i < pcap) pend[pbase + i] = pbuf[i];
    if (kHist) {
        for (uint32_t i = tid; i < GBH_EVEN_CLASSES * GBH_KA_MAX; i += GBH_CHECK_WG) {
            const uint32_t v = hl[i];
            if (v) atomicAdd(&out->hist[(i / GBH_KA_MAX) * GBH_NCAND + (i % GBH_KA_MAX)], v);
        }
    }
}

__global__ void __launch_bounds__(256) k_check_b(gbh_batch b, const uint32_t* __restrict__ coff, uint32_t ka,
                                                 uint32_t kend, uint32_t tail_t, const uint32_t* __restrict__ pending, uint32_t pcap,
                                                 gbh_segout* __restrict__ outs, uint64_t* __restrict__ ridge_more)
{
    const uint32_t s = blockIdx.y;
    const gbh_geom& g = b.geoms[s];
    if (g.n_gpu > g.seg_end) return;
    const uint32_t ridge_thresh = g.ridge_t;
    gbh_segout* out = outs + s;
    const uint64_t* map = seg_map(b, s);
    const uint32_t* pend = pending + (size_t)s * pcap;
    const uint32_t np = min(out->pending_n, pcap);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < np; x += gridDim.x * blockDim.x) {
        const uint32_t i = pend[x];
        const uint64_t n = g.seg_start + 2ull * i;
        const uint32_t cr = (uint32_t)(n % 30) >> 1;
        const ClassEvens e = class_evens(g, cr);
        const uint32_t k = (uint32_t)((n - e.n0) / 30);
        const uint32_t nj = min(c_cand_n[cr], kend);
        const uint32_t* ct = coff + ((size_t)s * 15 + cr) * GBH_NCAND;
        // ascending candidates, four probes in flight at a time; the first
        // set bit in candidate order is p(n)
        uint32_t jf = 0xFFFFFFFFu;
        for (uint32_t j = ka; j < nj && jf == 0xFFFFFFFFu; j += 4) {
            uint32_t sh[4];
            uint64_t wd[4];
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                wd[u] = 0;
                sh[u] = 0;
                if (j + u < nj) {
                    const uint32_t ent = ct[j + u];
                    const uint32_t c = ent >> 16, t = k + (ent & 0xFFFF);
                    sh[u] = t & 63;
                    wd[u] = map[(size_t)c * b.cw + (t >> 6)];
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u)
                if (jf == 0xFFFFFFFFu && j + u < nj && ((wd[u] >> sh[u]) & 1ull)) jf = j + u;
        }
        if (jf != 0xFFFFFFFFu) {
            atomicAdd(&out->hist[cr * GBH_NCAND + jf], 1u);
            const uint32_t pj = c_cand_p[cr * GBH_NCAND + jf];
            if (pj >= tail_t) atomicMin(&out->first[c_cand_pi[cr * GBH_NCAND + jf]], i);
            if (pj >= ridge_thresh) {
                const uint32_t sl = atomicAdd(&out->ridge_n, 1u);
                if (sl < GBH_RIDGE_CAP) {
                    out->ridge[2 * sl] = n;
                    out->ridge[2 * sl + 1] = pj;
                } else if (ridge_more && sl - GBH_RIDGE_CAP < GBH_RIDGE_MORE) {
                    uint64_t* m = ridge_more + 2 * ((size_t)s * GBH_RIDGE_MORE + (sl - GBH_RIDGE_CAP));
                    m[0] = n;
                    m[1] = pj;
                }
            }
            atomicAdd(&out->acct[1], 1u);
        } else {
            const uint32_t sl = atomicAdd(&out->left_n, 1u);
            if (sl < GBH_LEFT_CAP) out->left[sl] = i;
        }
    }
}

// ---------------------------------------------------------------- counter
__device__ __forceinline__ uint64_t range_mask(uint32_t x, uint32_t lo, uint32_t hi)
{
    const uint32_t b0 = x * 64, b1 = b0 + 64;
    if (hi <= b0 || lo >= b1 || lo >= hi) return 0;
    const uint32_t a = lo > b0 ? lo - b0 : 0, e = hi < b1 ? hi - b0 : 64;
    const uint64_t upto = e == 64 ? ~0ull : ((1ull << e) - 1);
    return upto & ~((1ull << a) - 1);
}

// Wave-wide minimum / maximum (every lane gets the result).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, o), hi = (uint32_t)__shfl_xor((int)(v >> 32), o);
        const uint64_t w = ((uint64_t)hi << 32) | lo;
        v = w > v ? w : v;
    }
    return v;
}

// The work-group's smallest first prime position, largest last one and
// largest gap key into the segment's outputs: one global atomic each per
// work-group (one per lane, all on the same three words, serialised at the
// cache and cost more than the walk itself). first = ~0u: no prime; last
// counts only with a prime; key 0: no gap.
__device__ __forceinline__ void flush_extremes(uint32_t first, uint32_t last, bool any, uint64_t key,
                                               gbh_segout* out, uint32_t* s_first, uint32_t* s_last,
                                               unsigned long long* s_key)
{
    const uint32_t tid = threadIdx.x;
    first = wave_min_u32(first);
    last = ~wave_min_u32(any ? ~last : 0xFFFFFFFFu);   // the largest last among lanes with a prime
    const bool wany = __ballot(any) != 0;
    key = wave_max_u64(key);
    if ((tid & 63) == 0) {
        if (wany) {
            atomicMin(s_first, first);
            atomicMax(s_last, last);
            atomicOr(s_last + 1, 1u);
        }
        if (key) atomicMax(s_key, (unsigned long long)key);
    }
    __syncthreads();
    if (tid == 0) {
        if (s_last[1]) {
            atomicMin(&out->pfirst, *s_first);
            atomicMax(&out->plast, *s_last);
        }
        if (*s_key) atomicMax((unsigned long long*)&out->gmax, *s_key);
    }
}

// Eight class words of 32 blocks each into four 64-bit masks in position
// order: bit 8 j + k of m[g] is bit 8 g + j of r[k]. First an 8 x 8 bit
// transpose inside every byte column across the eight registers (three
// rounds of block swaps: 4 x 4, 2 x 2, 1 x 1), after which byte g of r[j]
// holds bit 8 g + j of every class; then a 4 x 4 byte transpose of each
// half (byte permutes, {s0, s1} with s1 the low dword) gathers byte g of
// the eight registers into m[g].
__device__ __forceinline__ void class_words_to_blocks(uint32_t (&r)[8], uint64_t (&m)[4])
{
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t t = ((r[k] >> 4) ^ r[k + 4]) & 0x0F0F0F0Fu;
        r[k + 4] ^= t;
        r[k] ^= t << 4;
    }
#pragma unroll
    for (int k = 0; k < 8; k += (k & 1) ? 3 : 1) {   // 0, 1, 4, 5
        const uint32_t t = ((r[k] >> 2) ^ r[k + 2]) & 0x33333333u;
        r[k + 2] ^= t;
        r[k] ^= t << 2;
    }
#pragma unroll
    for (int k = 0; k < 8; k += 2) {
        const uint32_t t = ((r[k] >> 1) ^ r[k + 1]) & 0x55555555u;
        r[k + 1] ^= t;
        r[k] ^= t << 1;
    }
    uint32_t w[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t* v = r + 4 * h;
        const uint32_t t0 = __builtin_amdgcn_perm(v[1], v[0], 0x06020400u), t1 = __builtin_amdgcn_perm(v[1], v[0], 0x07030501u);
        const uint32_t t2 = __builtin_amdgcn_perm(v[3], v[2], 0x06020400u), t3 = __builtin_amdgcn_perm(v[3], v[2], 0x07030501u);
        w[h][0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
        w[h][1] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
        w[h][2] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
        w[h][3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) m[g] = ((uint64_t)w[1][g] << 32) | w[0][g];
}

__device__ __forceinline__ void note_gap(uint32_t g, uint32_t up, uint32_t* hl, gbh_segout* out, uint64_t& best)
{
    const uint32_t bin = g >> 1;
    atomicAdd(&hl[bin < GBH_GAP_BINS ? bin : GBH_GAP_BINS], 1u);
    if (g >= GBH_GAP_TAIL) {
        const uint32_t sl = atomicAdd(&out->gtail_n, 1u);
        if (sl < GBH_GAP_TCAP) { out->gtail[2 * sl] = up; out->gtail[2 * sl + 1] = g; }
    }
    const uint64_t key = ((uint64_t)g << 32) | (uint32_t)~up;
    if (key > best) best = key;
}

__global__ void __launch_bounds__(256) k_count_a(gbh_batch b, const gbh_countseg* __restrict__ css,
                                                 uint32_t* __restrict__ scratch, uint32_t swords,
                                                 gbh_segout* __restrict__ outs)
{
    constexpr uint32_t kSub = GBH_COUNT_SUB, kBits = 64 / kSub;   // pieces per word, blocks per piece
    __shared__ uint32_t hl[GBH_GAP_BINS + 1];
    __shared__ uint32_t cs[11], s_first, s_last[2];
    __shared__ unsigned long long s_key;
    const uint32_t tid = threadIdx.x, s = blockIdx.y;
    const gbh_countseg& a = css[s];
    const uint32_t npieces = a.nwords * kSub;
    if (!a.on || blockIdx.x * 256 >= npieces) return;
    // the walk's phase offsets, read once: pos0[k] / 2 (below 15) in 4-bit
    // field k
    uint32_t ph4 = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) ph4 |= (a.pos0[k] >> 1) << (4 * k);
    gbh_segout* out = outs + s;
    const uint64_t* map = seg_map(b, s);
    uint32_t* fl = scratch + (size_t)s * swords;
    for (uint32_t i = tid; i <= GBH_GAP_BINS; i += 256) hl[i] = 0;
    if (tid < 11) cs[tid] = 0;
    if (tid == 0) {
        s_first = 0xFFFFFFFFu;
        s_last[0] = s_last[1] = 0;
        s_key = 0;
    }
    __syncthreads();
    uint64_t best = 0;
    uint32_t t_first = 0xFFFFFFFFu, t_last = 0;
    bool t_any = false;
    uint32_t pcs[11];                       // primes per class, twins per pair: reduced per wave below
#pragma unroll
    for (int k = 0; k < 11; ++k) pcs[k] = 0;
    // pieces y, y + grid, ...: piece y = word y / kSub, blocks
    // [kBits (y % kSub), + kBits) of it (the kSub lanes of a word read the
    // same words); the work-group's histogram and counters are flushed once
    for (uint32_t y = blockIdx.x * 256 + tid; y < npieces; y += gridDim.x * 256) {
        const uint32_t x = y / kSub, sh = kBits * (y % kSub);
        const uint64_t pm = ((1ull << kBits) - 1) << sh;   // the piece's bits of the word
        uint32_t h[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint64_t w = map[(size_t)a.cls[k] * b.cw + x] & range_mask(x, a.lo[k], a.hi[k]) & pm;
            pcs[k] += (uint32_t)__popcll(w);
            h[k] = (uint32_t)(w >> sh);
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint64_t lw = map[(size_t)a.tk[j] * b.cw + x] & range_mask(x, a.tlo[j], a.thi[j]) & pm;
            const uint64_t* p2 = map + (size_t)a.tk2[j] * b.cw + x;
            const uint64_t pw = a.td[j] ? ((p2[0] >> 1) | (p2[1] << 63)) : p2[0];
            pcs[8 + j] += (uint32_t)__popcll(lw & pw);
        }
        // the walk: primes in position order (block t, then ascending
        // phase): the piece as four masks of eight blocks each, one step
        // per prime; a lane moves to its next non-empty mask by itself, so
        // the wave steps as often as its lane with the most primes
        uint64_t mk[4];
        class_words_to_blocks(h, mk);
        uint32_t first = 0xFFFFFFFFu, prev = 0;
        bool have = false;
        uint64_t m = mk[0];
        uint32_t g = 0;
        for (;;) {
            while (!m && g < 3) {
                ++g;
                m = g == 1 ? mk[1] : g == 2 ? mk[2] : mk[3];
            }
            if (!m) break;
            const uint32_t i = (uint32_t)__ffsll((long long)m) - 1;
            m &= m - 1;
            const uint32_t pos = 30u * (x * 64 + sh + 8 * g + (i >> 3)) + 2 * ((ph4 >> (4 * (i & 7))) & 15u);
            if (have) note_gap(pos - prev, pos, hl, out, best);
            else { first = pos; have = true; }
            prev = pos;
        }
        fl[2 * y] = first;
        fl[2 * y + 1] = have ? prev : 0xFFFFFFFFu;
        if (have) {
            t_first = min(t_first, first);
            t_last = max(t_last, prev);
            t_any = true;
        }
    }
#pragma unroll
    for (int k = 0; k < 11; ++k) {
        const uint32_t t = wave_total63(pcs[k]);   // the wave's total in lane 63
        if ((tid & 63) == 63 && t) atomicAdd(&cs[k], t);
    }
    flush_extremes(t_first, t_last, t_any, best, out, &s_first, s_last, &s_key);   // synchronises
    for (uint32_t i = tid; i <= GBH_GAP_BINS; i += 256)
        if (hl[i]) atomicAdd(&out->ghist[i], hl[i]);
    if (tid < 8 && cs[tid]) atomicAdd(&out->cnt[a.cls[tid]], cs[tid]);
    if (tid >= 8 && tid < 11 && cs[tid]) atomicAdd(&out->twins[tid - 8], cs[tid]);
}

// Gap from the last prime of piece y to the first prime of the next piece
// that has one (the walk across piece and word boundaries).
__global__ void __launch_bounds__(256) k_count_b(const gbh_countseg* __restrict__ css,
                                                 const uint32_t* __restrict__ scratch, uint32_t swords,
                                                 gbh_segout* __restrict__ outs)
{
    __shared__ uint32_t hl[GBH_GAP_BINS + 1];
    __shared__ uint32_t s_first, s_last[2];
    __shared__ unsigned long long s_key;
    const uint32_t tid = threadIdx.x, s = blockIdx.y;
    const gbh_countseg& a = css[s];
    const uint32_t npieces = a.nwords * GBH_COUNT_SUB;
    if (!a.on || blockIdx.x * 256 >= npieces) return;
    gbh_segout* out = outs + s;
    const uint32_t* fl = scratch + (size_t)s * swords;
    for (uint32_t i = tid; i <= GBH_GAP_BINS; i += 256) hl[i] = 0;
    if (tid == 0) {
        s_first = 0xFFFFFFFFu;
        s_last[0] = s_last[1] = 0;
        s_key = 0;
    }
    __syncthreads();
    uint64_t best = 0;
    for (uint32_t y = blockIdx.x * 256 + tid; y < npieces; y += gridDim.x * 256) {
        const uint32_t last = fl[2 * y + 1];
        if (last != 0xFFFFFFFFu) {
            for (uint32_t z = y + 1; z < npieces; ++z) {
                const uint32_t f = fl[2 * z];
                if (f != 0xFFFFFFFFu) { note_gap(f - last, f, hl, out, best); break; }
            }
        }
    }
    flush_extremes(0xFFFFFFFFu, 0, false, best, out, &s_first, s_last, &s_key);   // synchronises
    for (uint32_t i = tid; i <= GBH_GAP_BINS; i += 256)
        if (hl[i]) atomicAdd(&out->ghist[i], hl[i]);
}

// Words of class arrays that differ between two maps of one segment (bits
// below nc[c] only).
__global__ void __launch_bounds__(256) k_mapcmp(const uint64_t* __restrict__ a, const uint64_t* __restrict__ r,
                                                const gbh_geom* __restrict__ g, uint32_t cw, uint32_t* __restrict__ diff)
{
    const uint32_t c = blockIdx.y;
    const uint32_t nc = g->nc[c];
    const uint32_t nw = (nc + 63) / 64;
    uint32_t bad = 0;
    for (uint32_t x = blockIdx.x * 256 + threadIdx.x; x < nw; x += gridDim.x * 256) {
        uint64_t d = a[(size_t)c * cw + x] ^ r[(size_t)c * cw + x];
        if (x == nw - 1 && nc % 64) d &= (1ull << (nc % 64)) - 1;
        bad += d != 0;
    }
    bad = wave_sum(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(diff, bad);
}

__global__ void __launch_bounds__(256) k_fill(uint32_t* __restrict__ p, size_t n, uint32_t v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Output blocks of a batch reset for a new batch: zeros, first[] and pfirst ~0u.
__global__ void __launch_bounds__(256) k_reset_out(gbh_segout* __restrict__ outs)
{
    gbh_segout* o = outs + blockIdx.y;
    uint32_t* w = reinterpret_cast<uint32_t*>(o);
    const size_t n = offsetof(gbh_segout, left) / 4;                // the leftover list needs no reset
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t byte = i * 4;
        const bool ones = (byte >= offsetof(gbh_segout, first) && byte < offsetof(gbh_segout, ridge_n)) ||
                          byte == offsetof(gbh_segout, pfirst);
        w[i] = ones ? 0xFFFFFFFFu : 0u;
    }
}

}  // namespace gbh_dev
