"""Pins of the CPU oracle (oracle/gb_oracle.c) before anything is compared
with it: the reference's own stats_ref output on the golden windows, its
published p(n) table, the base-2 pseudoprimes below 2^32, and the
statistics-line restatement against the same fixtures."""
import pathlib
import subprocess

import pytest

import compare
from conftest import GOLDEN

FAST = ["w1e6", "w1e12s", "w4e18s", "w2p63s", "wgap1476", "wgap1442", "wtop"]


def _oracle(*args, stdin=None):
    return subprocess.run([str(compare.ORACLE), *map(str, args)], check=True, capture_output=True, text=True,
                          input=stdin).stdout


@pytest.mark.parametrize("name", FAST)
def test_oracle_matches_reference_stats_ref(name):
    """The restatement prints exactly what the reference's stats_ref printed
    (every field: p(n) histogram, first occurrences, primes / twins by
    residue, gap moments, tail and full gap histogram)."""
    a, b, ref = compare.load_fixture(GOLDEN / f"ref_{name}.txt")
    assert compare.oracle_ref(a, b) == ref


@pytest.mark.slow
@pytest.mark.parametrize("name", ["w1e12", "w2p63"])
def test_oracle_matches_reference_larger(name):
    a, b, ref = compare.load_fixture(GOLDEN / f"ref_{name}.txt")
    assert compare.oracle_ref(a, b) == ref


def test_miller_rabin_pseudoprimes():
    """All 10403 base-2 Fermat pseudoprimes below 2^32 (the reference's
    tests/data file) are composite to the oracle's Miller-Rabin; brute force
    agrees on every n below 2^17; and known primes near 2^64 are prime."""
    nums = [int(l) for l in (GOLDEN / "psp2_below_2_32.txt").read_text().splitlines() if l and l[0] != "#"]
    assert len(nums) == 10403
    assert all(pow(2, n - 1, n) == 1 for n in nums)                  # the file is what it claims
    out = _oracle("isprime", "-", stdin="\n".join(map(str, nums)) + "\n").split()
    assert out == ["0"] * len(nums)
    lim = 1 << 17
    sieve = bytearray([1]) * lim
    sieve[0] = sieve[1] = 0
    for i in range(2, int(lim ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(sieve[i * i::i]))
    out = _oracle("isprime", "-", stdin="\n".join(map(str, range(lim))) + "\n").split()
    assert [int(x) for x in out] == list(sieve)
    big = [18446744073709551557, 18446744073709551533, 9223372036854775783, 4611686018427387847]
    assert _oracle("isprime", "-", stdin="\n".join(map(str, big)) + "\n").split() == ["1"] * 4
    # strong pseudoprime to the bases 2..23 (149491 * 747451 * 34233211): needs 29, 31, 37
    assert _oracle("isprime", "-", stdin="3825123056546413051\n").split() == ["0"]


def test_published_pn_table():
    """p(S(p)) for the reference's published S(p) pairs (stats_ref.cpp:35-42)
    and the two pn_exact pins of host_tests.cpp:687-689."""
    table = [(9781, 3325581707333960528), (9629, 2795935116574469638), (9341, 906030579562279642),
             (9161, 887123803077837868), (8941, 555274351556750822), (8933, 258549426916149682),
             (8501, 255329126688555994), (8443, 121005022304007026), (8647, 2725058671961297876),
             (9203, 1348113579429547486), (8387, 3878297017437646306), (7307, 370581106766909188),
             (7, 30), (19, 98), (23, 220), (31, 308)]
    for p, s in table:
        assert int(_oracle("pn", s)) == p, (p, s)
    assert int(_oracle("pn", 4000000000169717704)) == 4133
    assert int(_oracle("pn", 4000000000000000000)) == 113


@pytest.mark.parametrize("name", ["w1e6", "w4e18s", "wgap1476"])
def test_gbstats_restatement_agrees_with_reference(name):
    """The oracle's GBSTATS fields pass the reference's own comparison
    against stats_ref, sampled everywhere and at the default sampling."""
    a, b, ref = compare.load_fixture(GOLDEN / f"ref_{name}.txt")
    for seg, sample, gaps in ((500000, 1, "exact"), (300000, 256, "exact"), (400000, 2, "sampled")):
        line, recs = compare.oracle_gbstats(a, b, seg, sample, gaps)
        bad = compare.compare_ref("GBSTATS v2 version=x " + line, "\n".join(recs), ref)
        assert not bad, (seg, sample, gaps, bad)


def test_compare_catches_a_planted_mistake():
    """A plausible slip (one gap bin moved, one prime class swapped) is
    caught by the comparison, so an agreeing run means something."""
    a, b, ref = compare.load_fixture(GOLDEN / "ref_w1e6.txt")
    line, recs = compare.oracle_gbstats(a, b, 500000, 1, "exact")
    kvs = compare.kv(line)
    p30 = kvs["primes30"].split(",")
    p30[0], p30[1] = p30[1], p30[0]
    bad_line = line.replace("primes30=" + kvs["primes30"], "primes30=" + ",".join(p30))
    assert compare.compare_ref("GBSTATS v2 version=x " + bad_line, "\n".join(recs), ref)
    gs = kvs["gsamp"].split(",")
    i = next(k for k, x in enumerate(gs) if x and int(x) > 1)
    gs[i] = str(int(gs[i]) - 1)
    gs[i + 1] = str(int(gs[i + 1] or 0) + 1)
    bad_line = line.replace("gsamp=" + kvs["gsamp"], "gsamp=" + ",".join(gs))
    assert compare.compare_ref("GBSTATS v2 version=x " + bad_line, "\n".join(recs), ref)


def test_segmentation_rule():
    """Sampled evens follow the reference's segmentation: spans clamp to
    give at least two segments, the remainder merges into the last one."""
    a, b = 1000000, 1002000            # 1001 evens
    line, _ = compare.oracle_gbstats(a, b, 467500000, 256, "exact")
    # span clamped to 1000 integers: segments [1e6, 1000998], [1001000, 1002000] (merged tail)
    assert compare.kv(line)["evens_sampled"] == "500"
    line, _ = compare.oracle_gbstats(a, b, 600, 2, "exact")
    # span 600: segments of 300, 300, 300 evens, then the 101-even rest (the
    # remaining 800 > 600 is not merged); segments 0 and 2 are sampled
    assert compare.kv(line)["evens_sampled"] == str(300 + 300)


def _pn_small(n, isp):
    """p(n) for small n from a sieve: the smallest odd prime p with n - p prime."""
    p = 3
    while not (isp[p] and isp[n - p]):
        p += 2
    return p


def test_oracle_ridges_against_brute_force():
    """The ridges mode (top-K largest p(n), ties by larger n) on a window a
    plain sieve covers, against the definition (include/ridge.hpp:52-60)."""
    a, b, floor, k = 1_000_000, 1_100_000, 40, 30
    lim = b + 2
    isp = bytearray([1]) * (lim + 1)
    isp[0] = isp[1] = 0
    for i in range(2, int(lim ** 0.5) + 1):
        if isp[i]:
            isp[i * i::i] = bytearray(len(isp[i * i::i]))
    want = sorted(((_pn_small(n, isp), n) for n in range(a, b + 1, 2)), reverse=True)
    want = [(n, p) for p, n in want if p >= floor][:k]
    got = [tuple(map(int, line.split())) for line in _oracle("ridges", a, b, floor, k, 4).splitlines()]
    assert got == want


def test_oracle_ridge_record_matches_reference_fixture():
    """The 4133 ridge near 4e18: the oracle's largest ridge of the window is
    the reference's pmax and first occurrence (tests/golden/ref_w4e18l.txt)."""
    fx = compare.kv((GOLDEN / "ref_w4e18l.txt").read_text())
    first = dict(item.split(":") for item in fx["first"].split(";"))
    top = _oracle("ridges", 4_000_000_000_169_000_000, 4_000_000_000_170_000_000, 3000, 5, 8).split()
    assert int(top[1]) == int(fx["pmax"]) == 4133
    assert int(top[0]) == int(first["4133"])
