"""The throughput slice's parity sample: 16 of the 2140 segments of
[4.5e18, 4.5e18 + 1e12 - 2], drawn with seed 0 (cfg3_segments.txt, written by
draw_configs.cpp), each verified as its own range and compared with the
reference's own stats_ref output over the same range (ref_cfg3_<k>.txt,
make_cfg3_fixtures.sh) by the reference's field comparison: once as plain
runs on one resident engine, once all 16 together through the batcher (whose
long batches take the direct harvest)."""
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


def _segments():
    out = []
    for line in (GOLDEN / "cfg3_segments.txt").read_text().splitlines():
        if line.startswith("#"):
            continue
        k, a, b = map(int, line.split())
        fa, fb, ref = compare.load_fixture(GOLDEN / f"ref_cfg3_{k}.txt")
        assert (fa, fb) == (a, b)
        out.append((k, a, b, ref))
    assert len(out) == 16
    return out


def test_slice_segments_match_stats_ref(stats_env):
    segs = _segments()
    hi = max(b for _, _, b, _ in segs)
    with gbamd.Engine(hi, SEGMENT) as eng:
        for k, a, b, ref in segs:
            for f in ("st.state", "st.log", "st.state.gaps"):
                (stats_env / f).unlink(missing_ok=True)
            _say(f"slice segment {k} [{a}, {b}]")
            r = eng.run(b, SEGMENT, a)
            assert r["leftovers"] == 0 and r["segments"] == 2
            bad = compare.compare_ref(r["stats_line"], (stats_env / "st.state").read_text(), ref)
            assert not bad, (k, bad)


def test_slice_segments_batched_match_stats_ref(stats_env):
    segs = _segments()
    lo, hi = min(a for _, a, _, _ in segs), max(b for _, _, b, _ in segs)
    _say(f"slice segments batched: {len(segs)} ranges")
    with gbamd.Engine(hi, SEGMENT) as eng, gbamd.Batcher(eng, hi, lo, SEGMENT) as bt:
        ids = [bt.submit(a, b) for _, a, b, _ in segs]
        assert sorted(bt.run_all()) == ids
        for (k, a, b, ref), i in zip(segs, ids):
            r = bt.result(i)
            bad = compare.compare_ref(r["stats_line"], r["state"], ref)
            assert not bad, (k, bad)
