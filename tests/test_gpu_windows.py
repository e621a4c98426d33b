"""Full-size GPU runs pinned against outside results.

* The exact range and call the bench times (one production range at 4e18:
  [4e18, 4e18 + 467,499,998], 233.75M evens, two half segments) against the
  reference's own stats_ref output of that range (tests/golden/ref_cfg2.txt).
* The published windows the reference validates its dense path on
  (tests/known_windows.py:45-53 of the reference, from an independent CUDA
  verifier and primecount): the prime count of each window, and the largest
  p(n) with the smallest n attaining it. These are many-segment runs of full
  467.5M segments in long batches, a shape no small fixture covers.
* The reference's end-to-end modes (tests/gpu_validate.sh:92-115 and
  tests/cli_tests.sh:44): statistics off, a run from n = 4, the --limit 4
  exit code, GB_BUCKET_VERIFY over long and ragged windows.
* The CPU fallback: phase B limited to its first candidate slots
  (GB_PHASEB_SLOTS, the role of the reference's GB_C30_PLAN override), so
  chosen evens become leftovers that the host resolves by Miller-Rabin; the
  statistics still equal the oracle's.
* The direct harvest's region overflow and its counted re-run (GB_HL_REGION).
"""
import subprocess
import sys

import pytest

import compare
import gbamd
from conftest import GOLDEN, PKG

pytestmark = pytest.mark.gpu

E18 = 4 * 10**18
CLI = str(PKG / "goldbach_amd")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _new_first(line: str) -> dict:
    v = compare.kv(line).get("new_first", "-")
    return {} if v in ("", "-") else {int(p): int(n) for p, n in (e.split(":") for e in v.split(";"))}


def test_bench_range_matches_stats_ref(stats_env):
    """The bench's own call, compared field by field with stats_ref."""
    a, b, ref = compare.load_fixture(GOLDEN / "ref_cfg2.txt")
    assert (a, b) == (E18, E18 + 467_499_998)
    _say(f"bench range [{a}, {b}]")
    r = gbamd.run_gpu(b, 467_500_000, start_even=a, collect_segments=True)
    assert r["segments"] == 2 and r["leftovers"] == 0
    segs = r["per_segment"]
    assert [(s["seg_start"], s["seg_end"]) for s in segs] == [(a, a + 233_749_998), (a + 233_750_000, b)]
    assert sum(s["evens"] for s in segs) == 233_750_000
    bad = compare.compare_ref(r["stats_line"], (stats_env / "st.state").read_text(), ref)
    assert not bad, bad


# name, S, N, primes in (S, N], largest p(n) (None: not published), smallest n with it
PUBLISHED = [
    ("4e18+1e10", E18, E18 + 10**10, 233_462_722, None, None),
    ("4e18+1e12", E18, E18 + 10**12, 23_346_564_662, 6073, E18 + 402_622_391_632),
    ("4e18+1e13", E18, E18 + 10**13, 233_465_239_142, 6421, E18 + 5_756_560_737_472),
    ("2^63", 9223371536854775808, 9223372536854775808, 22_899_962_096, 6661, 9223372067978243558),
]


@pytest.mark.parametrize("name,a,b,primes,pmax,n_at", PUBLISHED, ids=[w[0] for w in PUBLISHED])
def test_published_window(stats_env, name, a, b, primes, pmax, n_at):
    _say(f"published window {name} [{a}, {b}]")
    r = gbamd.run_gpu(b, 467_500_000, start_even=a)
    _say(f"  {r['segments']} segments, {r['evens_per_second'] / 1e9:.1f} G evens/s")
    assert r["leftovers"] == 0
    assert r["stats_primes"] == primes
    assert int(compare.kv(r["stats_line"])["primes"]) == primes
    if pmax is not None:
        assert r["stats_pmax"] == pmax
        assert _new_first(r["stats_line"])[pmax] == n_at


def _cli(*args, env=None):
    import os
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300, env=e)


def test_statistics_off_end_to_end(stats_env):
    """GB_STATS=0: a PASS without any statistics file."""
    _say("GB_STATS=0 [4e18, 4e18 + 1e10]")
    r = _cli("--start", str(E18), "--limit", str(E18 + 10**10), "--csv", env={"GB_STATS": "0"})
    assert r.returncode == 0, r.stderr
    assert "result=PASS" in r.stdout and "leftovers=0" in r.stdout
    assert not (stats_env / "st.state").exists() and not (stats_env / "st.log").exists()


def test_run_from_four_and_limit_four(stats_env):
    """From n = 4: the evens below the first checkable even go to the CPU
    (Miller-Rabin), the rest to the GPU; --limit 4 alone (4 = 2 + 2, no odd
    partition) exits 1, as the reference's CLI does."""
    _say("--start 4 --limit 2e6")
    r = gbamd.run_gpu(2_000_000, 500_000, start_even=4)
    assert r["segments"] == 4 and r["stats_on"] == 0
    c = _cli("--start", "4", "--limit", "3000000", "--segment", "1000000", "--csv")
    assert c.returncode == 0 and "result=PASS" in c.stdout, c.stderr
    assert _cli("--limit", "4").returncode == 1


@pytest.mark.parametrize("a,b,env", [
    (E18, E18 + 4 * 10**10, {}),
    (4999999999960000000, 5 * 10**18, {}),
    # lists across batch boundaries (5 segments a batch) with a ragged last
    # segment; a ragged first segment at the top
    (E18, 4000000051123456788, {"GB_BATCH": "5"}),
    (4999999948876543210, 5 * 10**18, {"GB_BATCH": "16"}),
], ids=["4e18+4e10", "5e18-4e10", "ragged-last", "ragged-first"])
def test_bucket_verify_long_windows(stats_env, monkeypatch, a, b, env):
    """Every segment's map rebuilt with every large prime on the wheel walk
    (no multiplier bitmap) and compared word for word on the GPU."""
    monkeypatch.setenv("GB_BUCKET_VERIFY", "1")
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    _say(f"bucket verify [{a}, {b}] {env}")
    r = gbamd.run_gpu(b, 467_500_000, start_even=a)
    assert r["leftovers"] == 0 and r["segments"] >= 2


@pytest.mark.parametrize("name,segment,env", [
    ("w1e12s", 500_000, {"GB_KA": "16", "GB_PHASEB_SLOTS": "24", "GB_STATS_SAMPLE": "1"}),
    ("w4e18s", 600_000, {"GB_KA": "32", "GB_PHASEB_SLOTS": "40", "GB_STATS_SAMPLE": "2"}),
    ("wgap1476", 1_000_000, {"GB_KA": "32", "GB_PHASEB_SLOTS": "40"}),
], ids=["1e12", "4e18", "gap1476"])
def test_cpu_fallback_leftovers(stats_env, monkeypatch, name, segment, env):
    """Phase B stops early: its unresolved evens are leftovers for the CPU
    route (cpu_min_goldbach_prime, statistics and ridges from the host), and
    the run's statistics still equal stats_ref's and the oracle's line."""
    a, b, ref = compare.load_fixture(GOLDEN / f"ref_{name}.txt")
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    _say(f"fallback case {name} [{a}, {b}] {env}")
    r = gbamd.run_gpu(b, segment, start_even=a)
    assert r["leftovers"] > 0
    bad = compare.compare_ref(r["stats_line"], (stats_env / "st.state").read_text(), ref)
    assert not bad, bad
    want, _ = compare.oracle_gbstats(a, b, segment, int(env.get("GB_STATS_SAMPLE", "256")), "exact")
    assert compare.strip_version(r["stats_line"]) == want


def test_phaseb_slots_must_exceed_ka(stats_env, monkeypatch):
    monkeypatch.setenv("GB_KA", "32")
    monkeypatch.setenv("GB_PHASEB_SLOTS", "32")
    with pytest.raises(gbamd.VerificationError, match="GB_PHASEB_SLOTS"):
        gbamd.run_gpu(2_000_000, 500_000, start_even=1_000_000)


def test_region_overflow_rerun_counted(stats_env, monkeypatch):
    """Direct-harvest regions of 32 entries overflow: every batch is re-run
    with counted lists, counted once, and the statistics equal stats_ref's."""
    a, b, ref = compare.load_fixture(GOLDEN / "ref_w4e18.txt")
    monkeypatch.setenv("GB_HL_MODE", "direct")
    monkeypatch.setenv("GB_HL_REGION", "32")
    _say(f"region overflow [{a}, {b}]")
    r = gbamd.run_gpu(b, 2_000_000, start_even=a)
    assert r["harvest_reruns"] >= 1 and r["leftovers"] == 0
    assert r["harvest_launches"] >= 1
    bad = compare.compare_ref(r["stats_line"], (stats_env / "st.state").read_text(), ref)
    assert not bad, bad


def test_ridge_records_beyond_the_output_block(stats_env):
    """A low ridge floor: a segment records far more than the 1024 ridges its
    output block holds; the rest come from the overflow buffer (65536 per
    segment in all, the reference's RIDGE_CAP), every record is considered
    (ridge_found above 1024 per segment), and the store's top 30 equals the
    oracle's 30 largest p(n) of the window."""
    a, b = E18, E18 + 2_000_000
    _say(f"ridge overflow [{a}, {b}] floor 700")
    c = _cli("--start", str(a), "--limit", str(b), "--segment", "1000000", "--csv",
             env={"GB_KA": "32", "GB_RIDGE_MIN": "700"})
    assert c.returncode == 0 and "result=PASS" in c.stdout, c.stderr
    row = compare.kv(c.stdout.strip().splitlines()[-1])
    assert int(row["segments"]) == 2 and int(row["ridge_min"]) == 700
    assert int(row["ridge_found"]) > 2 * 1024, row["ridge_found"]
    assert "more than" not in c.stderr          # no segment beyond the 65536 records
    rows = (stats_env / "ridges.csv").read_text().splitlines()
    got = [tuple(int(x) for x in r.split(",")[1:3]) for r in rows if not r.startswith("#")]
    out = subprocess.run([str(compare.ORACLE), "ridges", str(a), str(b), "700", "30", "8"], capture_output=True,
                         text=True, check=True).stdout
    assert got == [tuple(int(x) for x in line.split()) for line in out.splitlines()]


# the reference's --csv row: 35 keys in this order, then sysman= (and
# optional telemetry keys), src/main.cpp:192-226 of the reference
CSV_KEYS = ["version", "limit", "start", "segment", "split", "segments", "leftovers", "result", "init_s",
            "init_cache_s", "init_module_s", "init_upload_s", "gpu_s", "total_s", "sieve_s", "large_s", "check_s",
            "offsets_s", "table_s", "sync_s", "evens_per_sec", "tiles", "large_items", "ridge_min", "ridge_found",
            "ridge_dropped", "ridge_max", "stats", "stats_tail", "stats_sampled", "stats_pmax", "primes", "twins",
            "split_used", "sched"]


def test_csv_row_keys_in_reference_order(stats_env):
    _say("csv row [4e18, 4e18 + 2e6]")
    c = _cli("--start", str(E18), "--limit", str(E18 + 2_000_000), "--segment", "1000000", "--csv")
    assert c.returncode == 0, c.stderr
    row = c.stdout.strip().splitlines()[-1]
    keys = [w.split("=", 1)[0] for w in row.split()]
    assert keys[:len(CSV_KEYS)] == CSV_KEYS
    assert keys[len(CSV_KEYS)] == "sysman" and row.split()[len(CSV_KEYS)] == "sysman=off"
    v = compare.kv(row)
    assert v["result"] == "PASS" and v["sched"] == "driver" and int(v["segments"]) == 2
    assert int(v["primes"]) == 46640 and int(v["twins"]) == 1396   # the window's stats_ref figures
