"""Served verification: a request loop over one resident engine.

Requests (range verifications of any size) arrive over time; the loop admits
them by device memory, feeds every admitted request's segments to one
Batcher (continuous batching: a request that arrives while a batch runs
joins the next batch, and each step issues a batch before it collects the
previous one, so the host's collection overlaps the device's work), and
records each request's arrival, admission and completion times. The base
primes, mask tables and multiplier bitmap stay resident in the engine for the
whole loop. Every response's statistics are those of the request run alone
(the Batcher's contract).

Admission is by the device memory that requests occupy: the batch buffers.
The Batcher allocates them once, within the device budget (its segments per
batch are the most whose buffers fit); two batches are in flight at once, so
the buffers hold 2 x batch_segments segment slots, each slot costing the
allocation's bytes over its slots. An admitted request reserves slots for its
segments (all of them, or every slot for a request with more segments than
slots) until it completes; a request is admitted only when its reservation
fits beside those of the requests in flight (or nothing is in flight), so it
waits otherwise.

The reference has no serving mode (it verifies one range per process); the
request shape follows the served workload: start uniform over the even
numbers in [1e12, 5e18], length log-uniform in [1e4, 1e9] integers.
"""
from __future__ import annotations

import collections
import math
import time

import numpy as np

import gbamd

SEGMENT = 467_500_000
BUDGET = 64 << 30                       # device bytes the batch buffers may hold
REQ_START_LO, REQ_START_HI = 10**12, 5 * 10**18
REQ_LEN_LO, REQ_LEN_HI = 10**4, 10**9


def draw_requests(n: int, seed: int):
    """n requests [(start_even, limit)]: start uniform over the evens of
    [1e12, 5e18], length floor(exp(U(ln 1e4, ln 1e9))) integers, limit the
    last even of [start, start + length)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        start = 2 * int(rng.integers(REQ_START_LO // 2, REQ_START_HI // 2 + 1))
        length = int(math.floor(math.exp(rng.uniform(math.log(REQ_LEN_LO), math.log(REQ_LEN_HI)))))
        limit = (start + length - 1) & ~1
        out.append((start, max(limit, start + 2)))
    return out


def poisson_arrivals(n: int, rate: float, seed: int):
    """Arrival times (s from 0) of n requests at `rate` per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n)).tolist()


class Server:
    """The request loop over one engine, its batch buffers allocated within
    `budget` device bytes."""

    def __init__(self, engine: gbamd.Engine, max_limit: int, min_start: int, budget: int = BUDGET,
                 span: int = SEGMENT):
        self.engine = engine
        self.batcher = gbamd.Batcher(engine, max_limit, min_start, span, device_budget=budget)
        self.budget = budget
        self.span = span
        nbytes, self.slots = self.batcher.memory()
        self.slot_bytes = nbytes / self.slots
        self.peak_reserved = 0                # bytes (slots x slot bytes) reserved at once, at most

    def reservation(self, start: int, limit: int) -> int:
        """Segment slots a request reserves while it is admitted."""
        return min(self.batcher.segments_of(start, limit), self.slots)

    def serve(self, requests, arrivals, clock=time.perf_counter):
        """Runs the loop in real time over requests[i] arriving at
        arrivals[i] (seconds from the start). Returns, per request, a dict:
        arrival, admitted, completed (seconds), id, slots reserved and result."""
        n = len(requests)
        order = sorted(range(n), key=lambda i: arrivals[i])
        need = [self.reservation(a, b) for a, b in requests]
        t0 = clock()
        waiting = collections.deque()      # arrived, not admitted (no room)
        inflight = {}                      # batcher id -> request index
        used = 0                           # slots reserved by the requests in flight
        rec = [None] * n
        nxt = 0
        done = 0
        while done < n:
            now = clock() - t0
            while nxt < n and arrivals[order[nxt]] <= now:
                waiting.append(order[nxt])
                nxt += 1
            while waiting and (not inflight or used + need[waiting[0]] <= self.slots):
                i = waiting.popleft()
                a, b = requests[i]
                bid = self.batcher.submit(a, b)
                inflight[bid] = i
                used += need[i]
                self.peak_reserved = max(self.peak_reserved, int(used * self.slot_bytes))
                rec[i] = {"arrival": arrivals[i], "admitted": clock() - t0, "id": bid, "slots": need[i],
                          "reserved_after": used}
            fin, more = self.batcher.step()
            tnow = clock() - t0
            for bid in fin:
                i = inflight.pop(bid)
                used -= need[i]
                rec[i]["completed"] = tnow
                rec[i]["result"] = self.batcher.result(bid)
                done += 1
            if not more and not inflight and not waiting and nxt < n:
                wait = arrivals[order[nxt]] - (clock() - t0)
                if wait > 0:
                    time.sleep(wait)
        return rec

    def device_bytes(self) -> int:
        """Device bytes the batch buffers hold now."""
        return self.batcher.memory()[0]

    def close(self):
        self.batcher.close()


def fifo_latencies(arrivals, service):
    """Latencies of sequential execution in arrival order (one request at a
    time, each taking service[i] seconds)."""
    order = sorted(range(len(arrivals)), key=lambda i: arrivals[i])
    t = 0.0
    lat = [0.0] * len(arrivals)
    for i in order:
        t = max(t, arrivals[i]) + service[i]
        lat[i] = t - arrivals[i]
    return lat
