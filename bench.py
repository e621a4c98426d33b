"""Throughput benchmark of the MI355X verifier on one GPU.

Headline (value; one step): one production segment range at 4e18, every even
n in [4e18, 4e18 + 467,499,998] (233,750,000 evens), through the whole
pipeline with exact statistics and gaps, run as the reference runs that range
(two half segments by its at-least-two-segments rule). The engine (base
primes, tables, buffers) and the multiplier bitmap of the range are built
before the timed steps, as the reference builds its tables at start-up; each
step is one complete run (segmentation, GPU pipeline, CPU fallback,
statistics commit). value = evens verified per second over the timed steps.

The same JSON line carries the rest of the metric, each measured in this run:
  full_segments_4e18  [4e18, 4e18 + 1e12]: 2140 segments (2139 full and a
                      17.5M-integer tail), G evens/s, and
                      the window's published prime count, largest p(n) and
                      its first n checked against the run's statistics.
  batched             the 256 seeded uneven ranges of
                      tests/golden/cfg4_ranges.txt: plain per-range runs one
                      after another against one Batcher over all of them;
                      both rates, the ratio, and whether every range's
                      statistics (GBSTATS line and state text) are identical.
  serving             100 seeded requests arriving as Poisson streams of 10,
                      30, 100, 300 and 1000 requests/s into the request loop
                      of goldbach-l0_amd/serve.py: p50 / p99 latency and
                      throughput per rate, beside sequential plain execution
                      of the same requests (measured service times, FIFO).

roofline: for the kernel with the most device time per step (HIP events
around each launch on the engine's stream), its algorithmic HBM bytes per
launch (the primes, bitmap words and hit lists it must move once) over its
average launch time, against the MI355X HBM peak; traffic: HBM bytes per
launch from rocprofv3 FETCH_SIZE / WRITE_SIZE on this same workload, scaled
by the counters' reading on a plain device copy of known size. "binding"
names the resource that binds each main kernel, from the committed counter
summaries of this workload (profiles/*_counters.json): the busier of VALU
issue (busy VALU cycles of its waves) and LDS (cycles the LDS serves indexed
operations), each over the CU cycles of the kernel's live launch time.

cpu_baseline: the oracle's restatement of the reference's only CPU path for
these statistics (stats_ref: per-even ascending Miller-Rabin) over the whole
headline range, on the host cores this process may use (at most 16, the
host share of one GPU); its output is compared with the GPU run's statistics
(parity) and with the reference's own stats_ref output of the range
(tests/golden/ref_cfg2.txt).

Prints one JSON line. --gpus above 1 is refused: this build runs on one GPU.
--headline-only skips the sub-objects (for profiling), --no-cpu the CPU leg.
--copy-calibration runs only a 1 GiB device-to-device copy (three times), the
reference kernel for the traffic counters.

Other workloads (--workload; each prints its own JSON line):
  slice    the throughput slice: 1e12 consecutive integers from 4.5e18
           (2140 segments: 2139 full and a 17.5M-integer tail), one run per
           step; value = evens per second of wall time.
  batched  the batched sub-object above as a line of its own.
  serving  the serving sub-object above as a line of its own.
"""
import argparse
import ctypes
import json
import os
import pathlib
import sys
import tempfile
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "goldbach-l0_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

START = 4_000_000_000_000_000_000
LIMIT = START + 467_499_998
SEGMENT = 467_500_000
EVENS = (LIMIT - START) // 2 + 1
FULL_LIMIT = START + 10**12
# the window's published figures: primes in (4e18, 4e18 + 1e12], the largest
# smallest-Goldbach-prime p(n) and the first n attaining it
FULL_PUBLISHED = {"primes": 23_346_564_662, "pmax": 6073, "pmax_first_n": START + 402_622_391_632}
HBM_PEAK_GBS = 8000.0
CUS = 256
CLOCK_HZ = 2.4e9
PROFILES = ROOT / "profiles"
TRAFFIC_FILE = PROFILES / "k_harvest_4e18_segment_pair_traffic.json"
COUNTERS_SEGMENT = PROFILES / "pipeline_4e18_segment_pair_counters.json"
COUNTERS_SLICE = PROFILES / "pipeline_4p5e18_slice_counters.json"
CPU_THREADS_MAX = 16
SLICE_START = 4_500_000_000_000_000_000


def hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize.restype = ctypes.c_int
    return lib


def device_sync(lib):
    rc = lib.hipDeviceSynchronize()
    if rc:
        raise RuntimeError(f"hipDeviceSynchronize: {rc}")


def copy_calibration():
    lib = hip()
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)):
            raise RuntimeError("hipMalloc")
    if lib.hipMemset(a, 1, ctypes.c_size_t(n)):
        raise RuntimeError("hipMemset")
    for _ in range(3):
        if lib.hipMemcpy(b, a, ctypes.c_size_t(n), 3):     # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy")
    device_sync(lib)
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)


def say(**kv):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(json.dumps({"progress": kv}), file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(gpu_line, gpu_state):
    """The oracle's stats_ref restatement over the whole headline range, and
    its parity with the GPU run's statistics and with the golden fixture."""
    import compare

    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(CPU_THREADS_MAX, allowed))
    say(stage="cpu_baseline", evens=EVENS, threads=cores)
    t0 = time.perf_counter()
    line = compare.oracle_ref(START, LIMIT, cores)
    dt = time.perf_counter() - t0
    _, _, golden = compare.load_fixture(ROOT / "tests" / "golden" / "ref_cfg2.txt")
    bad = compare.compare_ref(gpu_line, gpu_state, line)
    bad_golden = compare.compare_ref(gpu_line, gpu_state, golden)
    return {"value": EVENS / dt / 1e9, "unit": "G evens/s", "cores": cores, "kind": "port",
            "sample": f"the whole headline range [{START}, {LIMIT}] ({EVENS} evens) through the stats_ref "
                      f"restatement (oracle/gb_oracle.c ref: per-even ascending Miller-Rabin, exact histogram, "
                      f"primes, twins, gaps), {dt:.1f} s wall on {cores} threads",
            "seconds": dt, "nproc": os.cpu_count(), "affinity_cpus": allowed, "cpu_model": cpu_model(),
            "parity": not bad, "parity_mismatches": bad[:4],
            "parity_reference_fixture": not bad_golden, "reference_fixture_mismatches": bad_golden[:4]}


def traffic():
    """HBM bytes per k_harvest launch from the committed counter summary."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
        return (t["fetch_kib_per_dispatch"] * 1024 / t["copy_fetch_factor"]
                + t["write_kib_per_dispatch"] * 1024 / t["copy_write_factor"])
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None


def binding(counter_file, kern, k):
    """The binding resource of k_harvest and k_window from the committed
    counters of the workload (per harvest launch; per batch for the window
    stage's two launches), over each kernel's live launch time in this run:
    VALU = SQ_ACTIVE_INST_VALU (quad-cycles of VALU work summed over waves;
    each CU's 4 SIMDs take one wave64 instruction per 4 cycles) over the CU
    cycles; LDS = SQ_LDS_IDX_ACTIVE over the CU cycles. The resource named is
    the busier one."""
    try:
        c = json.loads(counter_file.read_text())
    except (OSError, ValueError):
        return None
    out = {"counters": str(counter_file.relative_to(ROOT)), "cus": CUS, "clock_ghz": CLOCK_HZ / 1e9}
    for name in ("k_harvest", "k_window"):
        per = c.get("kernels", {}).get(name)
        sec, launches = kern[name][0], kern[name][1]
        if not per or not launches:
            continue
        launch_s = sec / launches
        cu_cycles = CUS * CLOCK_HZ * launch_s
        valu = per["SQ_ACTIVE_INST_VALU"] / cu_cycles
        lds = per["SQ_LDS_IDX_ACTIVE"] / cu_cycles
        out[name] = {"resource": "VALU issue" if valu >= lds else "LDS", "frac": max(valu, lds),
                     "valu_frac": valu, "lds_frac": lds,
                     "lds_bank_conflict_frac": per["SQ_LDS_BANK_CONFLICT"] / cu_cycles,
                     "valu_insts": per["SQ_INSTS_VALU"], "lds_insts": per["SQ_INSTS_LDS"],
                     "per": per.get("per", "launch"), "launch_ms": 1e3 * launch_s}
    return out


def stats_files(td):
    """Environment of throwaway statistics files under td."""
    return dict(GB_STATS_STATE=f"{td}/st.state", GB_STATS_LOG=f"{td}/st.log", GB_RECORDS_FILE=f"{td}/rec.txt",
                GB_STATS_LIVE=f"{td}/live.json", GB_RIDGE_FILE=f"{td}/ridges.csv", GB_RECORD_HOOK="")


def fresh(td):
    """A fresh campaign for the next run: its statistics files removed."""
    for f in ("st.state", "st.log", "st.state.gaps"):
        try:
            os.unlink(os.path.join(td, f))
        except FileNotFoundError:
            pass


def tmpfs():
    # the runs' statistics files are thrown away after each step: on tmpfs
    # their fsyncs measure memory, not whatever disk the host has
    return "/dev/shm" if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK) else None


def kernel_totals(runs):
    """per kernel: (seconds, launches, algorithmic bytes, bitmap words of those bytes)"""
    return {"k_harvest": (sum(r["harvest_kernel_seconds"] for r in runs), sum(r["harvest_launches"] for r in runs),
                          sum(r["harvest_bytes"] for r in runs), sum(r["harvest_words"] for r in runs)),
            "k_window": (sum(r["window_kernel_seconds"] for r in runs), sum(r["window_launches"] for r in runs),
                         sum(r["window_bytes"] for r in runs), 0.0)}


def hbm_roofline(kern):
    """The dominant kernel's HBM roofline over kernel totals."""
    dom = max(kern, key=lambda n: kern[n][0])
    sec, launches, nbytes, words = kern[dom]
    launch_s = sec / max(launches, 1)
    bytes_per_launch = nbytes / max(launches, 1)
    achieved = bytes_per_launch / launch_s / 1e9 if launch_s > 0 else 0.0
    # the same bytes with every bitmap word read counted as the 64-byte
    # sector the memory delivers for it (the words a prime reads are a
    # random access each: one sector per word, not 8 bytes)
    sector = bytes_per_launch + 56.0 * words / max(launches, 1)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "kernel": dom, "bytes_per_launch": bytes_per_launch,
            "bitmap_words_per_launch": words / max(launches, 1), "bytes_per_launch_sector": sector,
            "launch_ms": 1e3 * launch_s}


def per_step(runs, key):
    return sum(r[key] for r in runs) / len(runs)


def stage_ms(runs):
    return {"large": 1e3 * per_step(runs, "large_seconds"), "sieve": 1e3 * per_step(runs, "sieve_seconds"),
            "check": 1e3 * (per_step(runs, "verify_seconds") - per_step(runs, "count_seconds")),
            "count": 1e3 * per_step(runs, "count_seconds")}


def timed_runs(eng, td, lib, limit, start, steps, warmup):
    """warmup + steps runs of [start, limit] (fresh statistics each); the
    timed runs' results, their wall time, and the last run's state text."""
    for _ in range(warmup):
        fresh(td)
        eng.run(limit, SEGMENT, start)
    device_sync(lib)
    runs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        fresh(td)
        runs.append(eng.run(limit, SEGMENT, start))   # synchronous: every segment collected and committed
    device_sync(lib)
    wall = time.perf_counter() - t0
    return runs, wall, pathlib.Path(td, "st.state").read_text()


def measure_headline(args, lib):
    import gbamd

    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(LIMIT, SEGMENT)
        prep_s = eng.prepare(LIMIT, START)
        runs, wall, state = timed_runs(eng, td, lib, LIMIT, START, args.steps, args.warmup)
        init_s = eng.init_seconds
        eng.close()
    k = len(runs)
    kern = kernel_totals(runs)
    roof = hbm_roofline(kern)
    tr = traffic() if roof["kernel"] == "k_harvest" else None
    out = {
        "metric": "G evens/s at 4e18 w/ exact stats",
        "value": EVENS * k / wall / 1e9,
        "unit": "G evens/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / k,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: the integer range itself (no dataset)",
        "config": {"workload": "one production segment range at 4e18: evens in [4e18, 4e18 + 467,499,998], "
                               "233,750,000 evens, 2 half segments (the reference's at-least-two-segments rule), "
                               "exact statistics and gaps",
                   "start": START, "limit": LIMIT, "segment": SEGMENT, "segments": runs[-1]["segments"],
                   "stage_ms": stage_ms(runs),
                   "kernel_ms": {n: 1e3 * v[0] / k for n, v in kern.items()},
                   "gpu_loop_ms": 1e3 * per_step(runs, "gpu_seconds"),
                   "host_plan_ms": 1e3 * per_step(runs, "host_plan_seconds"),
                   "engine_init_s": init_s, "bitmap_prepare_s": prep_s},
        "roofline": {**{n: roof[n] for n in ("bound", "achieved", "peak", "unit", "frac")},
                     "traffic": tr,
                     **{n: roof[n] for n in ("kernel", "bytes_per_launch", "bitmap_words_per_launch",
                                             "bytes_per_launch_sector", "launch_ms")},
                     # counter traffic over the algorithmic bytes (8 bytes per bitmap word) and over the
                     # same bytes at the 64-byte sector of a random word read
                     "traffic_ratio": tr / roof["bytes_per_launch"] if tr and roof["bytes_per_launch"] else None,
                     "traffic_ratio_sector": (tr / roof["bytes_per_launch_sector"]
                                              if tr and roof["bytes_per_launch_sector"] else None),
                     "binding": binding(COUNTERS_SEGMENT, kern, k)},
    }
    return out, runs[-1]["stats_line"], state


def measure_full(lib, steps=2, warmup=1):
    """[4e18, 4e18 + 1e12] in full segments, checked against the published figures."""
    import compare
    import gbamd

    evens = (FULL_LIMIT - START) // 2 + 1
    say(stage="full_segments_4e18", limit=FULL_LIMIT)
    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(FULL_LIMIT, SEGMENT)
        prep_s = eng.prepare(FULL_LIMIT, START)
        runs, wall, _ = timed_runs(eng, td, lib, FULL_LIMIT, START, steps, warmup)
        eng.close()
    k = len(runs)
    last = runs[-1]
    first = compare.kv(last["stats_line"]).get("new_first", "-")
    firsts = {} if first in ("", "-") else {int(p): int(n) for p, n in (e.split(":") for e in first.split(";"))}
    got = {"primes": last["stats_primes"], "pmax": last["stats_pmax"],
           "pmax_first_n": firsts.get(last["stats_pmax"])}
    every = all((r["stats_primes"], r["stats_pmax"], r["leftovers"]) == (last["stats_primes"], last["stats_pmax"], 0)
                for r in runs)
    kern = kernel_totals(runs)
    return {"value": evens * k / wall / 1e9, "unit": "G evens/s", "start": START, "limit": FULL_LIMIT,
            "evens": evens, "segments": last["segments"], "steps": k, "warmup": warmup,
            "ms_per_step": 1e3 * wall / k, "published": FULL_PUBLISHED, "got": got,
            "check": "pass" if got == FULL_PUBLISHED and every else "fail",
            "stage_ms": stage_ms(runs), "kernel_ms": {n: 1e3 * v[0] / k for n, v in kern.items()},
            "bitmap_prepare_s": prep_s}


def load_cfg4():
    rows = [l.split() for l in (ROOT / "tests" / "golden" / "cfg4_ranges.txt").read_text().splitlines()
            if not l.startswith("#")]
    return [(int(a), int(b)) for a, b, _ in rows]


def measure_batched(lib, steps=3, warmup=1):
    """The 256 uneven ranges: plain per-range runs against one Batcher."""
    import gbamd

    ranges = load_cfg4()
    lo, hi = min(a for a, _ in ranges), max(b for _, b in ranges)
    evens = sum((b - a) // 2 + 1 for a, b in ranges)
    say(stage="batched", ranges=len(ranges))
    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(hi, SEGMENT)
        prep_s = eng.prepare(hi, lo)

        def plain(keep=False):
            got = []
            for a, b in ranges:
                fresh(td)
                r = eng.run(b, SEGMENT, a)
                if keep:
                    got.append((r["stats_line"], pathlib.Path(td, "st.state").read_text(), r["segments"],
                                r["leftovers"]))
            return got

        # the batcher and its batch buffers are set up once, before the timed
        # steps (as the engine is); each step submits every range and runs
        # them all
        bt = gbamd.Batcher(eng, hi, lo, SEGMENT)

        def batched(keep=False):
            ids = [bt.submit(a, b) for a, b in ranges]
            bt.run_all()
            got = []
            if keep:
                for i in ids:
                    r = bt.result(i)
                    got.append((r["stats_line"], r["state"], r["segments"], r["leftovers"]))
            return bt.totals(), bt.batch_segments, got

        # the untimed first pass also checks every range's statistics
        ref = plain(keep=True)
        _, _, mine = batched(keep=True)
        identical = sum(1 for x, y in zip(ref, mine) if x == y)
        for _ in range(warmup - 1):
            plain()
            batched()
        device_sync(lib)
        t_plain, t_batch, tot = [], [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            plain()
            device_sync(lib)
            t_plain.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            totals, kb, _ = batched()
            device_sync(lib)
            t_batch.append(time.perf_counter() - t0)
            tot.append(totals)
        bt.close()
        eng.close()
    k = len(t_batch)
    wall_b, wall_p = sum(t_batch) / k, sum(t_plain) / k
    kern = kernel_totals(tot)
    return {"value": evens / wall_b / 1e9, "unit": "G evens/s", "plain_value": evens / wall_p / 1e9,
            "ratio": wall_p / wall_b, "ranges": len(ranges), "evens": evens, "steps": k, "warmup": warmup,
            "plain_ms": 1e3 * wall_p, "batched_ms": 1e3 * wall_b, "batch_segments": kb,
            "identical_to_plain": identical, "parity": identical == len(ranges),
            "kernel_ms": {n: 1e3 * v[0] / k for n, v in kern.items()},
            "kernel_launches": {n: v[1] / k for n, v in kern.items()}, "bitmap_prepare_s": prep_s,
            # batches launched as one captured HIP graph per pass over the ranges (GB_GRAPH)
            "graph_launches": (tot[-1]["graph_launches"] - tot[0]["graph_launches"]) / (k - 1) if k > 1 else
                              tot[-1]["graph_launches"],
            "workload": "256 ranges, seed 1, starts log-uniform over [1e9, 5e18], lengths log-uniform over "
                        "[1e5, 4.675e8] integers (tests/golden/cfg4_ranges.txt)"}


def measure_serving():
    import gbamd
    import numpy as np
    import serve

    reqs = serve.draw_requests(100, 5)
    evens = sum((b - a) // 2 + 1 for a, b in reqs)
    max_limit = serve.REQ_START_HI + 2 * serve.REQ_LEN_HI
    rates = (10, 30, 100, 300, 1000)
    out = []
    say(stage="serving", requests=len(reqs))
    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(max_limit, serve.SEGMENT)
        prep_s = eng.prepare(max_limit, serve.REQ_START_LO)
        # sequential plain execution: each request's service time alone
        for a, b in reqs[:4]:
            fresh(td)
            eng.run(b, serve.SEGMENT, a)
        service = []
        for a, b in reqs:
            fresh(td)
            t0 = time.perf_counter()
            eng.run(b, serve.SEGMENT, a)
            service.append(time.perf_counter() - t0)
        warm = serve.Server(eng, max_limit, serve.REQ_START_LO)
        warm.serve(reqs[:8], [0.0] * 8)
        warm.close()
        for k, rate in enumerate(rates):
            arrivals = serve.poisson_arrivals(len(reqs), float(rate), 100 + k)
            srv = serve.Server(eng, max_limit, serve.REQ_START_LO)
            rec = srv.serve(reqs, arrivals)
            peak, held, slots = srv.peak_reserved, srv.device_bytes(), srv.slots
            srv.close()
            lat = np.array([r["completed"] - r["arrival"] for r in rec])
            span = max(r["completed"] for r in rec) - min(arrivals)
            seq = np.array(serve.fifo_latencies(arrivals, service))
            seq_span = max(a + l for a, l in zip(arrivals, seq)) - min(arrivals)
            out.append({"rate": rate, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                        "p99_ms": 1e3 * float(np.percentile(lat, 99)), "throughput_gevens_s": evens / span / 1e9,
                        "plain_p50_ms": 1e3 * float(np.percentile(seq, 50)),
                        "plain_p99_ms": 1e3 * float(np.percentile(seq, 99)),
                        "plain_throughput_gevens_s": evens / seq_span / 1e9,
                        "peak_reserved_gib": peak / 2**30, "batch_buffers_gib": held / 2**30,
                        "inflight_slots": slots})
            say(stage="serving", **out[-1])
        eng.close()
    return {"rates": out, "requests": len(reqs), "evens": evens, "budget_gib": serve.BUDGET / 2**30,
            "plain_service_ms_mean": 1e3 * sum(service) / len(service), "bitmap_prepare_s": prep_s,
            "workload": "serving: 100 requests (serve.draw_requests seed 5), Poisson arrivals at 10-1000 "
                        "requests/s, start uniform in [1e12, 5e18], length log-uniform in [1e4, 1e9] integers, "
                        "64 GiB device-memory admission budget"}


def bench_slice(args, lib):
    import gbamd

    start, limit = SLICE_START, SLICE_START + 10**12 - 2
    evens = (limit - start) // 2 + 1
    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(limit, SEGMENT)
        prep_s = eng.prepare(limit, start)
        runs, wall, _ = timed_runs(eng, td, lib, limit, start, args.steps, args.warmup)
        eng.close()
    k = len(runs)
    kern = kernel_totals(runs)
    roof = hbm_roofline(kern)
    print(json.dumps({
        "metric": "G evens/s at 4.5e18 w/ exact stats, 1e12-integer slice", "value": evens * k / wall / 1e9,
        "unit": "G evens/s", "n_gpus": 1, "steps": k, "warmup": args.warmup, "ms_per_step": 1e3 * wall / k,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: the integer range itself (no dataset)",
        "config": {"workload": "throughput slice: 1e12 consecutive integers from 4.5e18, 2140 segments (2139 full, "
                               "a 17.5M-integer tail), exact statistics and gaps",
                   "start": start, "limit": limit, "segment": SEGMENT, "segments": runs[-1]["segments"],
                   "gpu_loop_evens_per_s": per_step(runs, "evens_per_second"), "stage_ms": stage_ms(runs),
                   "kernel_ms": {n: 1e3 * v[0] / k for n, v in kern.items()},
                   "kernel_launches": {n: v[1] / k for n, v in kern.items()},
                   "host_plan_ms": 1e3 * per_step(runs, "host_plan_seconds"), "bitmap_prepare_s": prep_s},
        "roofline": {**roof, "binding": binding(COUNTERS_SLICE, kern, k)}}), flush=True)


def line_of(metric, value, unit, sub, steps, warmup, ms, data):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
            "data": data, "config": sub}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=("segment", "slice", "batched", "serving"), default="segment")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--headline-only", action="store_true", help="skip the full-segment, batched and serving parts")
    ap.add_argument("--copy-calibration", action="store_true", help="only the counter-calibration copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} refused: this build runs on one GPU (a multi-GPU campaign is disjoint "
              "ranges per GPU, independent runs merged on the host)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    steps_default, warm_default = {"segment": (20, 2), "slice": (2, 1), "batched": (3, 1), "serving": (1, 1)}[args.workload]
    args.steps = steps_default if args.steps is None else args.steps
    args.warmup = warm_default if args.warmup is None else args.warmup
    lib = hip()
    if args.workload == "slice":
        bench_slice(args, lib)
        return
    if args.workload == "batched":
        b = measure_batched(lib, args.steps, max(args.warmup, 1))
        print(json.dumps(line_of("batched vs plain per-range runs, G evens/s", b["value"], "G evens/s", b,
                                 b["steps"], b["warmup"], b["batched_ms"],
                                 "synthetic: 256 seeded ranges (tests/golden/cfg4_ranges.txt)")), flush=True)
        return
    if args.workload == "serving":
        s = measure_serving()
        top = s["rates"][-1]
        print(json.dumps(line_of("served p50/p99 latency (ms) and throughput", top["throughput_gevens_s"],
                                 "G evens/s at the highest rate", s, len(s["rates"]), 1, None,
                                 "synthetic: 100 seeded requests (goldbach-l0_amd/serve.py draw_requests, seed 5)")),
              flush=True)
        return
    out, gpu_line, gpu_state = measure_headline(args, lib)
    say(stage="headline", value=out["value"])
    if not args.headline_only:
        out["full_segments_4e18"] = measure_full(lib)
        out["batched"] = measure_batched(lib)
        out["serving"] = measure_serving()
    out["cpu_baseline"] = None if args.no_cpu else cpu_baseline(gpu_line, gpu_state)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
