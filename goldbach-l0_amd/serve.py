"""Served verification: a request loop over one resident engine.

Requests (range verifications of any size) arrive over time; the loop admits
them by device-memory footprint, feeds every admitted request's segments to
one Batcher (continuous batching: a request that arrives while a batch runs
joins the next batch, and each step issues a batch before it collects the
previous one, so the host's collection overlaps the device's work), and
records each request's arrival, admission and completion times. The base
primes, mask tables and multiplier bitmap stay resident in the engine for the
whole loop. Every response's statistics are those of the request run alone
(the Batcher's contract).

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


def footprint_bytes(start: int, limit: int, span: int = SEGMENT) -> int:
    """Device bytes a request's segments occupy while they are in a batch:
    per segment its class map (8 class arrays of about (evens + 4891) / 15
    bits), its hit lists (about 0.06 hits of 4 bytes per class bit at 4e18)
    and its pending list (1/32 of its evens plus 65536 entries of 4 bytes)."""
    evens = (limit - start) // 2 + 1
    seg_evens = max(1, min(evens // 2, span // 2))
    nseg = max(2, -(-evens // seg_evens))
    bits = (seg_evens + 4891) // 15 + 1          # per class array
    per_seg = bits + 2 * bits + (seg_evens // 32 + 65536) * 4
    return nseg * per_seg


class Server:
    """The request loop. `budget` bounds the summed footprint of admitted,
    unfinished requests (a request larger than the budget is admitted alone)."""

    def __init__(self, engine: gbamd.Engine, max_limit: int, min_start: int, budget: int = 64 << 30,
                 span: int = SEGMENT):
        self.engine = engine
        self.batcher = gbamd.Batcher(engine, max_limit, min_start, span)
        self.budget = budget
        self.span = span

    def serve(self, requests, arrivals, clock=time.perf_counter):
        """Runs the loop in real time over requests[i] arriving at
        arrivals[i] (seconds from the start). Returns, per request, a dict:
        arrival, admitted, completed (seconds), batch id and result."""
        n = len(requests)
        order = sorted(range(n), key=lambda i: arrivals[i])
        t0 = clock()
        waiting = collections.deque()      # arrived, not admitted (memory)
        inflight = {}                      # batcher id -> request index
        used = 0
        foot = [footprint_bytes(a, b, self.span) for a, b in requests]
        rec = [None] * n
        nxt = 0
        done = 0
        while done < n:
            now = clock() - t0
            while nxt < n and arrivals[order[nxt]] <= now:
                waiting.append(order[nxt])
                nxt += 1
            while waiting and (not inflight or used + foot[waiting[0]] <= self.budget):
                i = waiting.popleft()
                a, b = requests[i]
                bid = self.batcher.submit(a, b)
                inflight[bid] = i
                used += foot[i]
                rec[i] = {"arrival": arrivals[i], "admitted": clock() - t0, "id": bid}
            fin, more = self.batcher.step()
            tnow = clock() - t0
            for bid in fin:
                i = inflight.pop(bid)
                used -= foot[i]
                rec[i]["completed"] = tnow
                rec[i]["result"] = self.batcher.result(bid)
                done += 1
            if not more and not inflight and not waiting and nxt < n:
                wait = arrivals[order[nxt]] - (clock() - t0)
                if wait > 0:
                    time.sleep(wait)
        return rec

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
