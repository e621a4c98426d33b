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

}  // namespace gbh_dev
