"""GPU parity: the HIP pipeline (through the C ABI) against the CPU oracle
and the reference's golden stats_ref output.

* the class map of a segment equals Miller-Rabin primality of every q;
* a run's GBSTATS line and state file pass the reference's own comparison
  with stats_ref on the golden windows the reference validates on
  (tests/gpu_validate.sh:45-80 windows and segment sizes);
* on small windows the GBSTATS line (after its version field) and the state
  histograms are byte-identical to the oracle's restatement, at several
  sampling and gap settings.
"""
import subprocess
import sys

import numpy as np
import pytest

import compare
import gbamd
from conftest import GOLDEN, PKG

PKG_CLI = PKG / "goldbach_amd"

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _isprime_many(nums):
    out = subprocess.run([str(compare.ORACLE), "isprime", "-"], input="\n".join(map(str, nums)) + "\n",
                         capture_output=True, text=True, check=True).stdout.split()
    return np.array([int(x) for x in out], dtype=np.uint8)


@pytest.mark.parametrize("seg_start,limit,span", [
    (1_000_000_000_000, 1_000_002_000_000, 1_000_000),
    (4_000_000_000_000_000_000, 4_000_000_000_004_000_000, 2_000_000),
    (18446742974194923840, 18446742974197923840, 1_500_000),
    (10_000, 2_000_000, 1_000_000),
])
def test_segment_map_is_exact(seg_start, limit, span):
    _say(f"map case [{seg_start}, {limit}] span {span}")
    words, Q, nc = gbamd.segment_map(seg_start, limit, span)
    seg_end = limit if limit - seg_start <= span else seg_start + span - 2
    q_last = seg_end - 3
    for c in range(8):
        n = nc[c]
        bits = np.unpackbits(words[c].view(np.uint8), bitorder="little")[:n]
        q = [Q[c] + 30 * t for t in range(n)]
        assert q[-1] <= q_last
        want = _isprime_many(q)
        bad = np.nonzero(bits != want)[0]
        assert bad.size == 0, (c, [(int(t), Q[c] + 30 * int(t), int(bits[t]), int(want[t])) for t in bad[:5]])


# (fixture, --segment, environment) as the reference validates them
WINDOWS = [
    ("w1e6", 500_000, {}),
    ("w1e12s", 500_000, {"GB_STATS_SAMPLE": "1"}),
    ("w4e18s", 600_000, {}),
    ("w2p63s", 400_000, {"GB_STATS_SAMPLE": "2"}),
    ("wgap1476", 1_000_000, {}),
    ("wgap1476", 1_000_000, {"GB_GAPS": "sampled"}),
    ("wgap1442", 1_000_000, {}),
    ("wtop", 1_000_000, {}),
    ("w4e18", 2_000_000, {}),
    ("w44e18", 1_000_000, {"GB_STATS_SAMPLE": "1"}),
    ("w5e18end", 3_000_000, {}),
    ("w1e12", 1_000_000, {"GB_STATS_SAMPLE": "5"}),
    ("w2p63", 1_000_000, {}),
    ("wgapcross", 1_000_000, {}),
    ("w4e18l", 467_500_000, {}),
    ("w4e18l", 20_000_000, {"GB_STATS_SAMPLE": "3"}),
    ("cfg1", 467_500_000, {}),
    # windows far longer than the segments' class arrays (2^19 bits for
    # arrays of ~10^4): only a window's words that hold class bits are
    # sieved and stored
    ("w1e6", 500_000, {"GB_WIN_KB": "64"}),
    ("w4e18s", 600_000, {"GB_WIN_KB": "64", "GB_STATS_SAMPLE": "1"}),
    ("wtop", 1_000_000, {"GB_WIN_KB": "64"}),
    # the large-prime split: every large prime on the wheel walk (no
    # multiplier bitmap), and a small bitmap budget (wheel below ~1.7e7)
    ("w4e18", 2_000_000, {"GB_MT_GB": "0"}),
    # the harvest's forms: direct (one walk into work-group regions, then the
    # partition into window lists) and counted (two walks)
    ("w4e18", 2_000_000, {"GB_HL_MODE": "direct"}),
    ("w5e18end", 3_000_000, {"GB_HL_MODE": "direct", "GB_MT_GB": "8"}),
    ("wtop", 1_000_000, {"GB_HL_MODE": "direct"}),
    ("w4e18", 2_000_000, {"GB_HL_MODE": "counted"}),
    # one stream: the window stage's sieve phase after the harvest
    ("w4e18", 2_000_000, {"GB_OVERLAP": "0"}),
    # no band of wheel primes among the bitmap primes, and a wide one
    ("w4e18", 2_000_000, {"GB_HL_WHEEL_HITS": "0"}),
    ("w2p63s", 400_000, {"GB_HL_WHEEL_HITS": "64"}),
    ("w44e18", 1_000_000, {"GB_MT_GB": "8"}),
    # every segment's map rebuilt with all large primes on the wheel walk and
    # compared word for word (the reference's GB_BUCKET_VERIFY self-check)
    ("w4e18", 2_000_000, {"GB_BUCKET_VERIFY": "1"}),
    ("w5e18end", 3_000_000, {"GB_BUCKET_VERIFY": "2"}),
]


@pytest.mark.parametrize("name,segment,env", WINDOWS, ids=[f"{w[0]}-{w[1]}-{'-'.join(w[2].values()) or 'dflt'}" for w in WINDOWS])
def test_run_matches_reference_stats_ref(stats_env, monkeypatch, name, segment, env):
    a, b, ref = compare.load_fixture(GOLDEN / f"ref_{name}.txt")
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    _say(f"run case {name} [{a}, {b}] segment {segment} env {env}")
    r = gbamd.run_gpu(b, segment, start_even=a)
    assert r["leftovers"] == 0
    log = (stats_env / "st.log").read_text().strip().splitlines()[-1]
    assert log == r["stats_line"].strip()
    bad = compare.compare_ref(log, (stats_env / "st.state").read_text(), ref)
    assert not bad, bad


SMALL = [
    (1_000_000, 3_000_000, 500_000, "256", "exact"),
    (1_000_000, 3_000_000, 300_000, "2", "sampled"),
    (4_000_000_000_000_000_000, 4_000_000_000_002_000_000, 700_000, "1", "exact"),
    (1425172824437600000, 1425172824437800000, 60_000, "3", "sampled"),
    (1425172824437600000, 1425172824437800000, 60_000, "256", "0"),
    (9223372036853775808, 9223372036855775808, 250_000, "4", "exact"),
]


@pytest.mark.parametrize("a,b,seg,sample,gaps", SMALL)
def test_gbstats_line_byte_exact(stats_env, monkeypatch, a, b, seg, sample, gaps):
    monkeypatch.setenv("GB_STATS_SAMPLE", sample)
    monkeypatch.setenv("GB_GAPS", gaps)
    _say(f"gbstats case [{a}, {b}] segment {seg} sample {sample} gaps {gaps}")
    r = gbamd.run_gpu(b, seg, start_even=a)
    want_line, want_recs = compare.oracle_gbstats(a, b, seg, int(sample), gaps)
    got = compare.strip_version(r["stats_line"])
    if gaps == "0":      # no gap fields at all
        want_line = " ".join(t for t in want_line.split() if not t.startswith(("gaps=", "gd1=", "gmax=", "gtail=", "gwalk=", "gsmom=", "gsamp=")))
    assert got == want_line
    assert compare.state_hist_records((stats_env / "st.state").read_text()) == want_recs


def test_state_merge_and_overlap_refusal(stats_env):
    """A second run over the same range is not counted twice; an adjacent
    run merges into the campaign totals."""
    _say("state merge case")
    gbamd.run_gpu(2_000_000, 500_000, start_even=1_000_000)
    st1 = (stats_env / "st.state").read_text()
    gbamd.run_gpu(2_000_000, 500_000, start_even=1_000_000)
    assert (stats_env / "st.state").read_text() == st1
    gbamd.run_gpu(3_000_000, 500_000, start_even=2_000_002)
    lines = (stats_env / "st.state").read_text().splitlines()
    assert [l for l in lines if l.startswith("range ")] == ["range 1000000 3000000"]
    # 500001 + 500000 evens; segment 0 of each run sampled: 250000 + 250000
    assert [l for l in lines if l.startswith("evens ")] == ["evens 1000001 500000"]


def test_ridges_match_oracle(stats_env, monkeypatch):
    """Goldbach ridges (include/ridge.hpp): the store's top 30 after a run equals
    the oracle's 30 largest p(n) of the window (floor 1800, above every phase-A
    candidate), the store format is the reference's v2, and a second run over
    the same window keeps the store unchanged."""
    a, b = 4_000_000_000_169_000_000, 4_000_000_000_170_000_000
    monkeypatch.setenv("GB_RIDGE_MIN", "1800")
    _say(f"ridges [{a}, {b}]")
    r = gbamd.run_gpu(b, 300_000, start_even=a)
    rows = (stats_env / "ridges.csv").read_text().splitlines()
    assert rows[0].startswith("# goldbach-l0 ridge store v2")
    got = [tuple(int(x) for x in row.split(",")[1:3]) for row in rows if not row.startswith("#")]
    out = subprocess.run([str(compare.ORACLE), "ridges", str(a), str(b), "1800", "30", "8"], capture_output=True,
                         text=True, check=True).stdout
    want = [tuple(int(x) for x in line.split()) for line in out.splitlines()]
    assert got == want
    assert got[0] == (4_000_000_000_169_717_704, 4133)
    for row in rows:
        if not row.startswith("#"):
            rank, n, p, q, seg, found = (int(x) for x in row.split(","))
            assert q == n - p and a <= seg <= n
    before = (stats_env / "ridges.csv").read_text()
    (stats_env / "st.state").unlink()
    (stats_env / "st.log").unlink()
    gbamd.run_gpu(b, 300_000, start_even=a)
    assert (stats_env / "ridges.csv").read_text() == before


def _fnv1a_words(data: bytes) -> int:
    h = 0xcbf29ce484222325
    for (w,) in __import__("struct").iter_unpack("<I", data):
        h = ((h ^ w) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_prime_cache_gbp2(stats_env):
    """The reference's GBP2 base-prime cache (src/primes.cpp:125-284) through the
    CLI: written as <stem>.<n> with a valid header, FNV-1a payload hash and the
    odd primes <= n; reused as is; a corrupted cache is ignored and rewritten;
    a cache for a larger bound serves a smaller one by its prefix."""
    import struct
    cli = str(PKG_CLI)
    stem = stats_env / "pc"
    a, b = 1_000_000_000_000, 1_000_000_200_000
    n = int(b ** 0.5) + 16

    def run(lim, start=a):
        return subprocess.run([cli, "--start", str(start), "--limit", str(lim), "--prime-cache", str(stem)],
                              capture_output=True, text=True, timeout=120)
    _say("prime cache: first run")
    r = run(b)
    assert r.returncode == 0, r.stderr
    f = stats_env / f"pc.{n}"
    raw = f.read_bytes()
    magic, lim, cnt, h = raw[:4], *struct.unpack("<QQQ", raw[4:28])
    assert magic == b"GBP2" and lim == n and len(raw) == 28 + 4 * cnt
    primes = list(struct.unpack(f"<{cnt}I", raw[28:]))
    assert primes == compare.odd_primes(n)
    assert h == _fnv1a_words(raw[28:])
    mtime = f.stat().st_mtime_ns
    r = run(b)
    assert r.returncode == 0 and "prime cache" not in r.stderr and f.stat().st_mtime_ns == mtime
    bad = bytearray(raw)
    bad[100] ^= 1
    f.write_bytes(bytes(bad))
    r = run(b)
    assert r.returncode == 0 and "payload hash mismatch" in r.stderr
    assert f.read_bytes() == raw
    # a smaller bound: served by the prefix of pc.<n>, no new file
    a2 = 500_000_000_000
    r = run(a2 + 2_000, a2)
    n2 = int((a2 + 2_000) ** 0.5) + 16
    assert r.returncode == 0 and n2 < n and not (stats_env / f"pc.{n2}").exists()


def test_long_batches_direct_harvest_maps_verified(stats_env, monkeypatch):
    """Full 467.5M segments in batches of 8 (the direct harvest's production
    shape): every segment's class map is rebuilt with every large prime on
    the wheel walk and compared word for word on the GPU (GB_BUCKET_VERIFY),
    and the statistics equal the counted harvest's."""
    a = 4_500_000_000_000_000_000
    b = a + 8 * 467_500_000 - 2
    monkeypatch.setenv("GB_BUCKET_VERIFY", "1")
    _say(f"long direct case [{a}, {b}]")
    r = gbamd.run_gpu(b, 467_500_000, start_even=a)
    assert r["segments"] == 8 and r["leftovers"] == 0
    monkeypatch.delenv("GB_BUCKET_VERIFY")
    for f in ("st.state", "st.log", "st.state.gaps"):
        (stats_env / f).unlink(missing_ok=True)
    monkeypatch.setenv("GB_HL_MODE", "counted")
    _say("long counted case")
    r2 = gbamd.run_gpu(b, 467_500_000, start_even=a)
    assert compare.strip_version(r2["stats_line"]) == compare.strip_version(r["stats_line"])
    # the long runs' window extension (large wheel primes below 2^20 sieved
    # in the windows) off: the same statistics
    monkeypatch.delenv("GB_HL_MODE")
    monkeypatch.setenv("GB_WIN_P_LONG", "524288")
    state = (stats_env / "st.state").read_text()
    for f in ("st.state", "st.log", "st.state.gaps"):
        (stats_env / f).unlink(missing_ok=True)
    _say("long case without the window extension")
    r3 = gbamd.run_gpu(b, 467_500_000, start_even=a)
    assert compare.strip_version(r3["stats_line"]) == compare.strip_version(r["stats_line"])
    assert (stats_env / "st.state").read_text() == state
