"""Served verification (goldbach-l0_amd/serve.py over the Batcher C ABI):
requests arriving faster than they complete overlap in shared batches; every
response equals a plain run of that request alone (GBSTATS line and state
text byte for byte, segments and leftovers), and the 20 shortest pass the
reference's comparison with their stats_ref fixtures
(tests/golden/ref_serve_short.txt); the device budget bounds the device
memory the batch buffers take, and requests that do not fit wait.
"""
import ctypes
import sys

import pytest

import compare
import gbamd
import serve
from conftest import GOLDEN

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
        # the 20 shortest against the reference's stats_ref output
        short = {}
        for line in (GOLDEN / "ref_serve_short.txt").read_text().splitlines():
            a, b, ref = line.split(" ", 2)
            short[(int(a), int(b))] = ref
        assert len(short) == 20
        for (a, b), r in zip(reqs, rec):
            if (a, b) in short:
                bad = compare.compare_ref(r["result"]["stats_line"], r["result"]["state"], short[(a, b)])
                assert not bad, (a, b, bad)


def _free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_budget_bounds_device_memory_and_requests_wait(stats_env):
    """A 2 GiB budget: the batcher's buffers (measured by the device's free
    memory around its allocation, and by its own account) stay within it
    before and after serving; requests each reserving more than half the
    in-flight segment slots run one at a time (each admitted only after the
    one before completed), and each equals its run alone."""
    budget = 2 << 30
    with gbamd.Engine(MAX_LIMIT, serve.SEGMENT) as eng:
        eng.prepare(MAX_LIMIT, serve.REQ_START_LO)     # resident tables first: not part of the budget
        free0 = _free_device_bytes()
        srv = serve.Server(eng, MAX_LIMIT, serve.REQ_START_LO, budget=budget)
        held = srv.device_bytes()
        free1 = _free_device_bytes()
        _say(f"budget {budget}: {srv.slots} slots of {srv.slot_bytes:.0f} bytes, {held} held, "
             f"{free0 - free1} taken from the device")
        assert srv.slots >= 2
        assert held <= budget and free0 - free1 <= budget
        k = srv.slots // 2 + 1                 # segments per request: no two fit at once
        reqs = [(s, s + k * serve.SEGMENT - 2) for s in (10**15, 4 * 10**18, 10**17, 2 * 10**18)]
        assert all(srv.reservation(a, b) == k for a, b in reqs)
        rec = srv.serve(reqs, [0.0] * len(reqs))
        assert srv.peak_reserved <= held <= budget
        assert srv.device_bytes() <= budget and free0 - _free_device_bytes() <= budget
        srv.close()
        spans = sorted((r["admitted"], r["completed"]) for r in rec)
        for (a0, c0), (a1, c1) in zip(spans, spans[1:]):
            assert a1 >= c0, spans
        for (a, b), r in zip(reqs, rec):
            for f in ("st.state", "st.log", "st.state.gaps"):
                (stats_env / f).unlink(missing_ok=True)
            alone = eng.run(b, serve.SEGMENT, a)
            assert r["result"]["stats_line"] == alone["stats_line"], (a, b)


def test_budget_below_one_segment_refused(stats_env):
    with gbamd.Engine(MAX_LIMIT, serve.SEGMENT) as eng:
        with pytest.raises(gbamd.VerificationError, match="device budget"):
            serve.Server(eng, MAX_LIMIT, serve.REQ_START_LO, budget=1 << 20)
