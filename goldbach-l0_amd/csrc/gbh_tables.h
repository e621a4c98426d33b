// Compile-time tables of the device kernels (the mod-30 class structure and
// the mask-stage primes 7..61; reference goldbach.cl:45-120 constants and
// config.hpp:136-146 MASK_MAX_P).
#pragma once
#include <stdint.h>

namespace gbh_dev {

// residues of the eight class arrays
__device__ constexpr uint32_t kRes[8] = {1, 7, 11, 13, 17, 19, 23, 29};
// x^-1 mod 30 for units x (0 elsewhere)
__device__ constexpr uint32_t kInv30[30] = {0, 1, 0, 0, 0, 0, 0, 13, 0, 0, 0, 11, 0, 7, 0,
                                            0, 0, 23, 0, 19, 0, 0, 0, 17, 0, 0, 0, 0, 0, 29};
// mask-stage primes (every odd prime below 64 except 3 and 5, which the
// class arrays exclude)
__device__ constexpr uint32_t kSmall[15] = {7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61};

constexpr uint64_t pat_of(uint32_t p)
{
    uint64_t m = 0;
    for (uint32_t b = 0; b < 64; b += p) m |= 1ull << b;
    return m;
}
// bits 0, p, 2p, ... below 64
__device__ constexpr uint64_t kPat[15] = {pat_of(7),  pat_of(11), pat_of(13), pat_of(17), pat_of(19),
                                          pat_of(23), pat_of(29), pat_of(31), pat_of(37), pat_of(41),
                                          pat_of(43), pat_of(47), pat_of(53), pat_of(59), pat_of(61)};
// phase advance of a lane's next word (256 words = 16384 bits later), mod p
__device__ constexpr uint32_t kStep256[15] = {16384 % 7,  16384 % 11, 16384 % 13, 16384 % 17, 16384 % 19,
                                              16384 % 23, 16384 % 29, 16384 % 31, 16384 % 37, 16384 % 41,
                                              16384 % 43, 16384 % 47, 16384 % 53, 16384 % 59, 16384 % 61};

// class index of a unit residue mod 30 (0xFF elsewhere)
__device__ constexpr uint8_t c_cls_of[30] = {0xFF, 0, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 1, 0xFF, 0xFF,
                                             0xFF, 2, 0xFF, 3, 0xFF, 0xFF, 0xFF, 4, 0xFF, 5,
                                             0xFF, 0xFF, 0xFF, 6, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 7};

// The 210-wheel: the 48 residues coprime to 2 3 5 7.
constexpr bool cop210(uint32_t r) { return r % 2 && r % 3 && r % 5 && r % 7; }
struct Wheel210 {
    uint8_t next[210];   // distance from r to the next coprime residue (0: r coprime)
    uint8_t idx[210];    // rank of a coprime residue
    uint8_t inc[48];     // gap from the rank-i residue to the next (cyclic)
};
constexpr Wheel210 make_wheel210()
{
    Wheel210 w{};
    uint32_t res[48] = {};
    uint32_t n = 0;
    for (uint32_t r = 0; r < 210; ++r)
        if (cop210(r)) { w.idx[r] = uint8_t(n); res[n++] = r; }
    for (uint32_t r = 0; r < 210; ++r) {
        uint32_t d = 0;
        while (!cop210((r + d) % 210)) ++d;
        w.next[r] = uint8_t(d);
    }
    for (uint32_t i = 0; i < 48; ++i) w.inc[i] = uint8_t(i + 1 < 48 ? res[i + 1] - res[i] : res[0] + 210 - res[47]);
    return w;
}
constexpr Wheel210 kWheel = make_wheel210();
__device__ constexpr Wheel210 c_wheel = kWheel;
#define c_w210_next c_wheel.next
#define c_w210_idx c_wheel.idx
#define c_w210_inc c_wheel.inc

}  // namespace gbh_dev
