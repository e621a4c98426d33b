"""Served verification (goldbach-l0_amd/serve.py over the Batcher C ABI):
requests arriving faster than they complete overlap in shared batches, and
every response equals a plain run of that request alone (GBSTATS line and
state text byte for byte, segments and leftovers)."""
import sys

import pytest

import gbamd
import serve

pytestmark = pytest.mark.gpu

MAX_LIMIT = serve.REQ_START_HI + 2 * serve.REQ_LEN_HI


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_served_responses_equal_runs_alone(stats_env):
    reqs = serve.draw_requests(24, 11)
    arrivals = serve.poisson_arrivals(len(reqs), 2000.0, 11)
    _say(f"serving case: {len(reqs)} requests at 2000/s")
    with gbamd.Engine(MAX_LIMIT, serve.SEGMENT) as eng:
        srv = serve.Server(eng, MAX_LIMIT, serve.REQ_START_LO)
        rec = srv.serve(reqs, arrivals)
        srv.close()
        # continuous batching: requests were admitted while others were in flight
        overlapped = sum(1 for i, r in enumerate(rec)
                         if any(o["admitted"] <= r["admitted"] < o["completed"] for j, o in enumerate(rec) if j != i))
        assert overlapped >= len(reqs) // 2, overlapped
        for (a, b), r in zip(reqs, rec):
            assert r["arrival"] <= r["admitted"] <= r["completed"]
            for f in ("st.state", "st.log", "st.state.gaps"):
                (stats_env / f).unlink(missing_ok=True)
            alone = eng.run(b, serve.SEGMENT, a)
            got = r["result"]
            assert got["stats_line"] == alone["stats_line"], (a, b)
            assert got["state"] == (stats_env / "st.state").read_text(), (a, b)
            assert (got["segments"], got["leftovers"]) == (alone["segments"], alone["leftovers"]), (a, b)


def test_admission_respects_the_memory_budget(stats_env):
    """With a budget below any two requests' footprints, requests run one at
    a time (each admitted only after the one before completed)."""
    reqs = serve.draw_requests(6, 12)
    arrivals = [0.0] * len(reqs)
    budget = min(serve.footprint_bytes(a, b) for a, b in reqs)
    with gbamd.Engine(MAX_LIMIT, serve.SEGMENT) as eng:
        srv = serve.Server(eng, MAX_LIMIT, serve.REQ_START_LO, budget=budget)
        rec = srv.serve(reqs, arrivals)
        srv.close()
    spans = sorted((r["admitted"], r["completed"]) for r in rec)
    for (a0, c0), (a1, c1) in zip(spans, spans[1:]):
        assert a1 >= c0, spans
