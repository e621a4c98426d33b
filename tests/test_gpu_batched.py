"""Batched ranges through one pipeline (gbamd.Batcher, through the C ABI).

* Golden windows of every magnitude (1e6 .. 2^64 - 2^40) submitted together
  and run as shared batches: each range passes the reference's comparison
  with its stats_ref fixture (tests/stats_compare.py:21-88 restated in
  oracle/compare.py).
* The 256 seeded ranges of the batched workload (cfg4_ranges.txt): each
  range's GBSTATS line and state text are byte-identical to a plain run of
  that range alone (the plain path is pinned to the oracle by
  test_gpu_parity.py), and the 32 shortest pass the comparison with their
  stats_ref fixtures.
"""
import sys

import pytest

import compare
import gbamd
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

SEGMENT = 467_500_000


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cfg4():
    rows = [l.split() for l in (GOLDEN / "cfg4_ranges.txt").read_text().splitlines() if not l.startswith("#")]
    return [(int(a), int(b)) for a, b, _ in rows]


GOLD = ["w1e6", "w1e12s", "w4e18s", "w2p63s", "wgap1476", "wgap1442", "wtop", "w44e18", "w5e18end", "w1e12", "w2p63",
        "wgapcross", "w4e18"]


def test_batched_golden_windows_match_stats_ref(stats_env):
    fx = [compare.load_fixture(GOLDEN / f"ref_{n}.txt") for n in GOLD]
    lo, hi = min(a for a, _, _ in fx), max(b for _, b, _ in fx)
    span = 1_000_000
    _say(f"batched golden windows: {len(fx)} ranges in [{lo}, {hi}], segment {span}")
    with gbamd.Engine(hi, span) as eng, gbamd.Batcher(eng, hi, lo, span) as bt:
        ids = [bt.submit(a, b) for a, b, _ in fx]
        done = bt.run_all()
        assert sorted(done) == ids
        for name, (a, b, ref), i in zip(GOLD, fx, ids):
            r = bt.result(i)
            assert r["leftovers"] == 0, name
            assert r["segments"] >= 2, name
            bad = compare.compare_ref(r["stats_line"], r["state"], ref)
            assert not bad, (name, bad)


def test_batched_ranges_equal_plain_runs(stats_env):
    ranges = _cfg4()
    lo, hi = min(a for a, _ in ranges), max(b for _, b in ranges)
    _say(f"batched workload: {len(ranges)} ranges in [{lo}, {hi}]")
    with gbamd.Engine(hi, SEGMENT) as eng:
        eng.prepare(hi, lo)
        with gbamd.Batcher(eng, hi, lo, SEGMENT) as bt:
            ids = [bt.submit(a, b) for a, b in ranges]
            done = bt.run_all()
            assert sorted(done) == ids
            got = [bt.result(i) for i in ids]
        short = {}
        for line in (GOLDEN / "ref_cfg4_short.txt").read_text().splitlines():
            a, b, ref = line.split(" ", 2)
            short[(int(a), int(b))] = ref
        assert len(short) == 32
        for k, (a, b) in enumerate(ranges):
            for f in ("st.state", "st.log", "st.state.gaps"):
                (stats_env / f).unlink(missing_ok=True)
            r = eng.run(b, SEGMENT, a)
            g = got[k]
            assert g["stats_line"] == r["stats_line"], (k, a, b)
            assert g["state"] == (stats_env / "st.state").read_text(), (k, a, b)
            assert (g["segments"], g["leftovers"]) == (r["segments"], r["leftovers"]), (k, a, b)
            if (a, b) in short:
                bad = compare.compare_ref(g["stats_line"], g["state"], short.pop((a, b)))
                assert not bad, (k, a, b, bad)
        assert not short


def test_batcher_refuses_ranges_outside_its_bounds(stats_env):
    with gbamd.Engine(10_000_000_000, 1_000_000) as eng, gbamd.Batcher(eng, 5_000_000_000, 1_000_000, 1_000_000) as bt:
        with pytest.raises(gbamd.VerificationError, match="above the batcher"):
            bt.submit(1_000_000, 6_000_000_000)
        with pytest.raises(gbamd.VerificationError, match="below the batcher"):
            bt.submit(10_000, 2_000_000)
        i = bt.submit(2_000_000, 3_000_000)
        with pytest.raises(gbamd.VerificationError, match="not finished"):
            bt.result(i)
        assert bt.run_all() == [i]
        assert bt.result(i)["stats_on"]
