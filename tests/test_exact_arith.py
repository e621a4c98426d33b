"""The kernels' loop-free quotient corrections, restated in numpy with the
device's float formats, against exact integer arithmetic.

* Window sieve (`first_off_big`): d < 2^31, 67 <= p < 2^21, a float32
  quotient d * rcp(p), the residue's own float32 quotient, then one step each
  way, must give d mod p exactly for a reciprocal within one ulp of 1/p.
* Harvest (`bm_setup`, `bm_single`): floor(qb0 / p) from a double quotient
  with the reciprocal refined by one Newton step from a float32 seed is
  within one of the integer quotient for bitmap primes (p^3 > q, q < 2^64),
  and (`div_big`) from a double reciprocal for wheel primes p >= 2^16.
Rounding of float32 products and casts follows IEEE round-to-nearest, as on
the device; the hardware reciprocal is modelled as 1/p rounded and moved by
-1, 0 or +1 ulp."""
import numpy as np

RNG = np.random.default_rng(20261203)


def _rcp32(p, k):
    r = (np.float32(1.0) / p.astype(np.float32)).astype(np.float32)
    up = np.nextafter(r, np.float32(np.inf))
    dn = np.nextafter(r, np.float32(0))
    return np.where(k > 0, up, np.where(k < 0, dn, r)).astype(np.float32)


def _first_off_residue(d, p, rp, second=True):
    """d mod p by the window kernel's two float32 quotients and one step each way."""
    q1 = (d.astype(np.float32) * rp).astype(np.float32).astype(np.uint32)   # truncating cast
    r = (d - q1 * p).astype(np.int64)
    r = np.where(r >= 2**31, r - 2**32, r)                                    # the int32 view
    if second:
        q2 = (r.astype(np.float32) * rp).astype(np.float32).astype(np.int64)  # truncates toward zero
        r = r - q2 * p.astype(np.int64)
    r = np.where(r < 0, r + p, np.where(r >= p, r - p, r))
    return r


def test_window_remainder_random():
    n = 2_000_000
    p = RNG.integers(67, 1 << 21, n, dtype=np.uint64).astype(np.uint32)
    p[::2] = RNG.integers(67, 2000, n // 2, dtype=np.uint64).astype(np.uint32)
    d = RNG.integers(0, 1 << 31, n, dtype=np.uint64).astype(np.uint32)
    for k in (-1, 0, 1):
        rp = _rcp32(p, np.full(n, k))
        got = _first_off_residue(d, p, rp)
        assert np.array_equal(got, (d % p).astype(np.int64)), k
    # the second quotient is needed: one float quotient and one step each
    # way leave small primes' residues wrong
    one = _first_off_residue(d, p, _rcp32(p, np.zeros(n)), second=False)
    assert not np.array_equal(one, (d % p).astype(np.int64))


def test_window_remainder_edges():
    ps = np.array([67, 71, 73, 79, 97, 4093, 4099, 131071, 524269, 524287, 1048573, 2097143], dtype=np.uint32)
    d_hi = np.arange((1 << 31) - 200_000, 1 << 31, dtype=np.uint64).astype(np.uint32)
    d_lo = np.arange(0, 200_000, dtype=np.uint32)
    for p0 in ps:
        for d in (d_hi, d_lo):
            p = np.full(d.shape, p0, dtype=np.uint32)
            for k in (-1, 0, 1):
                got = _first_off_residue(d, p, _rcp32(p, np.full(d.shape, k)))
                assert np.array_equal(got, (d % p).astype(np.int64)), (p0, k)


def _quotient_error(q, p, inv):
    """floor(float64(q) * inv) minus the exact quotient (python ints)."""
    est = np.floor(q.astype(np.float64) * inv)
    return [int(e) - int(a) // int(b) for e, a, b in zip(est, q, p)]


def test_bitmap_quotient_within_one():
    n = 200_000
    q = RNG.integers(1 << 40, (1 << 64) - 1, n, dtype=np.uint64, endpoint=True)
    cube = np.cbrt(q.astype(np.float64))
    p = np.minimum(cube * RNG.uniform(1.0, 1000.0, n), 2**32 - 1).astype(np.uint64)
    p = np.maximum(p, (cube + 2).astype(np.uint64))            # bitmap primes: p^3 > q
    for k in (-1, 0, 1):
        inv0 = _rcp32(p.astype(np.uint32), np.full(n, k)).astype(np.float64)
        pd = p.astype(np.float64)
        inv = inv0 + inv0 * (1.0 - pd * inv0)                     # one Newton step in double
        err = _quotient_error(q, p, inv)
        assert max(abs(e) for e in err) <= 1, k


def test_wheel_quotient_within_one():
    n = 200_000
    q = RNG.integers(1 << 40, (1 << 64) - 1, n, dtype=np.uint64, endpoint=True)
    p = RNG.integers(1 << 16, 1 << 32, n, dtype=np.uint64)
    p[:1000] = 1 << 16
    pd = p.astype(np.float64)
    r = 1.0 / pd
    inv = r + r * (1.0 - pd * r)
    err = _quotient_error(q, p, inv)
    assert max(abs(e) for e in err) <= 1
