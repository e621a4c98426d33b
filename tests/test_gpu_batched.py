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
* Graph replays: one window run again and again with GB_GRAPH=1 (the
  engine's graphs re-captured and updated in place, run after run),
  between eager runs of the same and of another window, every run checked
  against the window's stats_ref fixture; a batcher with and without graphs
  gives identical results.
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


def test_graph_replays_match_stats_ref(stats_env, monkeypatch):
    a, b, ref = compare.load_fixture(GOLDEN / "ref_w4e18s.txt")
    a2, b2, ref2 = compare.load_fixture(GOLDEN / "ref_w1e12s.txt")
    span = 700_000

    def run(lo, hi, graph):
        monkeypatch.setenv("GB_GRAPH", "1" if graph else "0")
        for f in ("st.state", "st.log", "st.state.gaps"):
            (stats_env / f).unlink(missing_ok=True)
        r = eng.run(hi, span, lo)
        return r, (stats_env / "st.state").read_text()

    with gbamd.Engine(b, span) as eng:
        plan = [(a, b, ref, True), (a, b, ref, True), (a, b, ref, False), (a, b, ref, True), (a2, b2, ref2, True),
                (a, b, ref, True), (a2, b2, ref2, False), (a2, b2, ref2, True)]
        for k, (lo, hi, rf, graph) in enumerate(plan):
            _say(f"graph replay {k}: [{lo}, {hi}] graph={graph}")
            r, state = run(lo, hi, graph)
            assert (r["graph_launches"] > 0) == graph, k
            assert r["leftovers"] == 0, k
            bad = compare.compare_ref(r["stats_line"], state, rf)
            assert not bad, (k, bad)


def test_batcher_with_and_without_graphs_agree(stats_env, monkeypatch):
    ranges = sorted(_cfg4(), key=lambda x: x[1] - x[0])[:48]
    lo, hi = min(x for x, _ in ranges), max(y for _, y in ranges)
    out = {}
    with gbamd.Engine(hi, SEGMENT) as eng:
        for graph in ("1", "0", "1"):
            monkeypatch.setenv("GB_GRAPH", graph)
            with gbamd.Batcher(eng, hi, lo, SEGMENT) as bt:
                ids = [bt.submit(x, y) for x, y in ranges]
                assert sorted(bt.run_all()) == ids
                res = [(bt.result(i)["stats_line"], bt.result(i)["state"]) for i in ids]
                launches = bt.totals()["graph_launches"]
            _say(f"batcher GB_GRAPH={graph}: {launches} graph launches")
            assert (launches > 0) == (graph == "1")
            if graph in out:
                assert out[graph] == res
            out[graph] = res
        assert out["1"] == out["0"]
