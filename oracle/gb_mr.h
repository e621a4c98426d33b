/* Deterministic 64-bit Miller-Rabin for the oracle (test infrastructure).
 *
 * Restates the reference's is_prime_u64 (src/primes.cpp:445-480): trial
 * division by the twelve primes 2..37 (n equal to one of them is prime),
 * then strong-probable-prime tests to the same twelve bases, which is exact
 * for every n < 3.3e24 and so for all of uint64_t. The modular arithmetic
 * here is Montgomery multiplication (R = 2^64) instead of the reference's
 * 128-bit remainder; the verdict is the same function of n.
 */
#ifndef GB_MR_H
#define GB_MR_H

#include <stdint.h>

typedef unsigned __int128 gb_u128;

/* -n^-1 mod 2^64 for odd n (Newton iteration on the 2-adic inverse). */
static inline uint64_t gb_mont_ninv(uint64_t n)
{
    uint64_t x = n;                     /* correct to 3 bits for odd n */
    for (int i = 0; i < 5; ++i) x *= 2 - n * x;
    return (uint64_t)0 - x;
}

/* a * b * 2^-64 mod n, for a, b < n, n odd. */
static inline uint64_t gb_mont_mul(uint64_t a, uint64_t b, uint64_t n, uint64_t ninv)
{
    const gb_u128 t = (gb_u128)a * b;
    const uint64_t m = (uint64_t)t * ninv;
    const gb_u128 u = t + (gb_u128)m * n;        /* may carry past 2^128 */
    uint64_t hi = (uint64_t)(u >> 64);
    const int carry = u < t;
    if (carry || hi >= n) hi -= n;
    return hi;
}

static inline int gb_is_prime_u64(uint64_t n)
{
    static const uint64_t small[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i)
        if (n % small[i] == 0) return n == small[i];
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    const uint64_t ninv = gb_mont_ninv(n);
    const uint64_t r1 = (uint64_t)(((gb_u128)1 << 64) % n);        /* 1 in Montgomery form */
    const uint64_t r2 = (uint64_t)(((gb_u128)r1 * r1) % n);        /* R^2 mod n */
    const uint64_t mone = n - r1;                                   /* -1 in Montgomery form */
    for (int i = 0; i < 12; ++i) {
        const uint64_t a = gb_mont_mul(small[i] % n, r2, n, ninv);  /* base to Montgomery form */
        uint64_t x = r1, b = a, e = d;
        while (e) {
            if (e & 1) x = gb_mont_mul(x, b, n, ninv);
            b = gb_mont_mul(b, b, n, ninv);
            e >>= 1;
        }
        if (x == r1 || x == mone) continue;
        int comp = 1;
        for (int k = 1; k < s; ++k) {
            x = gb_mont_mul(x, x, n, ninv);
            if (x == mone) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

#endif
