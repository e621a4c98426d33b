"""Throughput benchmark of the MI355X verifier on one GPU.

Workload (one step): one production segment range at 4e18, every even n in
[4e18, 4e18 + 467,499,998] (233,750,000 evens), through the whole pipeline
with exact statistics and gaps, run as the reference runs that range (two
half segments by its at-least-two-segments rule). The engine (base primes,
tables, buffers) and the multiplier bitmap of the range are built before the
timed steps, as the reference builds its tables at start-up; each step is one
complete run (segmentation, GPU pipeline, CPU fallback, statistics commit).
value = evens verified per second over the timed steps.

roofline: the kernel with the most device time per step (HIP events around
each launch on the engine's stream), its algorithmic HBM bytes per launch
(the primes, bitmap words and hit lists it must move once) over its average
launch time, against the MI355X HBM peak; traffic: HBM bytes per launch
measured by rocprofv3 FETCH_SIZE / WRITE_SIZE on this same command, scaled by
the counters' reading on a plain device copy of known size (the committed
profile summary), or null without one.

cpu_baseline: the oracle's restatement of the reference's only CPU path for
these statistics (stats_ref: per-even ascending Miller-Rabin) on a bounded
sample of the same range, on the host's cores (at most 16).

Prints one JSON line. --gpus above 1 is refused: this build runs on one GPU.
--copy-calibration runs only a 1 GiB device-to-device copy (three times), the
reference kernel for the traffic counters.

Other workloads (--workload; each prints its own JSON line):
  slice    the throughput slice: 1e12 consecutive integers from 4.5e18
           (2140 segments: 2139 full and a 17.5M-integer tail), one run per
           step; value = evens per second of wall time.
  batched  the 256 seeded ranges of tests/golden/cfg4_ranges.txt (starts
           log-uniform over [1e9, 5e18], lengths log-uniform over
           [1e5, 4.675e8]): plain per-range runs one after another against
           one Batcher over all of them (every range's statistics identical
           either way); value = batched evens per second, with the plain rate
           and the ratio.
  serving  100 seeded requests (start uniform in [1e12, 5e18], length
           log-uniform in [1e4, 1e9]) arriving as Poisson streams of 10, 30,
           100, 300 and 1000 requests/s into the request loop of
           goldbach-l0_amd/serve.py (continuous batching, 64 GB admission
           budget, resident tables): p50 / p99 latency and throughput per
           rate, beside sequential plain execution of the same requests
           (measured service times, FIFO).
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
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "k_harvest_4e18_segment_pair_traffic.json"
CPU_SAMPLE_PER_CORE = 2_000_000
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


def cpu_baseline():
    """The oracle's stats_ref restatement on a bounded sample of the range."""
    import compare

    cores = min(16, os.cpu_count() or 1)
    n_evens = CPU_SAMPLE_PER_CORE * cores
    a = START
    b = START + 2 * (n_evens - 1)
    t0 = time.perf_counter()
    compare.oracle_ref(a, b, cores)
    dt = time.perf_counter() - t0
    return {"value": n_evens / dt / 1e9, "unit": "G evens/s", "cores": cores, "kind": "port",
            "sample": f"stats_ref restatement (oracle/gb_oracle.c ref: per-even Miller-Rabin, exact histogram, "
                      f"primes, twins, gaps) over [{a}, {b}]: {n_evens} evens of the same range, {dt:.1f} s wall "
                      f"on {cores} threads"}


def traffic():
    """HBM bytes per k_harvest launch from the committed counter summary."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
        return (t["fetch_kib_per_dispatch"] * 1024 / t["copy_fetch_factor"]
                + t["write_kib_per_dispatch"] * 1024 / t["copy_write_factor"])
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None


def stats_files(td):
    """Environment of throwaway statistics files under td."""
    return dict(GB_STATS_STATE=f"{td}/st.state", GB_STATS_LOG=f"{td}/st.log", GB_RECORDS_FILE=f"{td}/rec.txt",
                GB_STATS_LIVE=f"{td}/live.json", GB_RIDGE_FILE=f"{td}/ridges.csv", GB_RECORD_HOOK="")


def fresh(td):
    for f in ("st.state", "st.log", "st.state.gaps"):
        pathlib.Path(td, f).unlink(missing_ok=True)


def tmpfs():
    # the runs' statistics files are thrown away after each step: on tmpfs
    # their fsyncs measure memory, not whatever disk the host has
    return "/dev/shm" if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK) else None


def harvest_roofline(runs):
    """The dominant kernel's roofline over a list of run results."""
    kern = {"k_harvest": (sum(r["harvest_kernel_seconds"] for r in runs), sum(r["harvest_launches"] for r in runs),
                          sum(r["harvest_bytes"] for r in runs)),
            "k_window": (sum(r["window_kernel_seconds"] for r in runs), sum(r["window_launches"] for r in runs),
                         sum(r["window_bytes"] for r in runs))}
    dom = max(kern, key=lambda n: kern[n][0])
    sec, launches, nbytes = kern[dom]
    launch_s = sec / max(launches, 1)
    bytes_per_launch = nbytes / max(launches, 1)
    achieved = bytes_per_launch / launch_s / 1e9 if launch_s > 0 else 0.0
    return kern, {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "kernel": dom, "bytes_per_launch": bytes_per_launch,
                  "launch_ms": 1e3 * launch_s}


def bench_slice(args):
    import gbamd

    start, limit = SLICE_START, SLICE_START + 10**12 - 2
    evens = (limit - start) // 2 + 1
    lib = hip()
    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(limit, SEGMENT)
        prep_s = eng.prepare(limit, start)
        for _ in range(args.warmup):
            fresh(td)
            eng.run(limit, SEGMENT, start)
        device_sync(lib)
        runs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fresh(td)
            runs.append(eng.run(limit, SEGMENT, start))
        device_sync(lib)
        wall = time.perf_counter() - t0
        eng.close()
    k = len(runs)
    kern, roof = harvest_roofline(runs)
    per_step = lambda key: sum(r[key] for r in runs) / k
    print(json.dumps({
        "metric": "G evens/s at 4.5e18 w/ exact stats, 1e12-integer slice", "value": evens * k / wall / 1e9,
        "unit": "G evens/s", "n_gpus": 1, "steps": k, "warmup": args.warmup, "ms_per_step": 1e3 * wall / k,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: the integer range itself (no dataset)",
        "config": {"workload": "throughput slice: 1e12 consecutive integers from 4.5e18, 2140 segments (2139 full, "
                               "a 17.5M-integer tail), exact statistics and gaps",
                   "start": start, "limit": limit, "segment": SEGMENT, "segments": runs[-1]["segments"],
                   "gpu_loop_evens_per_s": per_step("evens_per_second"),
                   "stage_ms": {"large": 1e3 * per_step("large_seconds"), "sieve": 1e3 * per_step("sieve_seconds"),
                                "check": 1e3 * (per_step("verify_seconds") - per_step("count_seconds")),
                                "count": 1e3 * per_step("count_seconds")},
                   "kernel_ms": {n: 1e3 * v[0] / k for n, v in kern.items()},
                   "kernel_launches": {n: v[1] / k for n, v in kern.items()},
                   "host_plan_ms": 1e3 * per_step("host_plan_seconds"), "bitmap_prepare_s": prep_s},
        "roofline": roof}), flush=True)


def load_cfg4():
    rows = [l.split() for l in (ROOT / "tests" / "golden" / "cfg4_ranges.txt").read_text().splitlines()
            if not l.startswith("#")]
    return [(int(a), int(b)) for a, b, _ in rows]


def bench_batched(args):
    import gbamd

    ranges = load_cfg4()
    lo, hi = min(a for a, _ in ranges), max(b for _, b in ranges)
    evens = sum((b - a) // 2 + 1 for a, b in ranges)
    lib = hip()
    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(hi, SEGMENT)
        prep_s = eng.prepare(hi, lo)

        def plain():
            for a, b in ranges:
                fresh(td)
                eng.run(b, SEGMENT, a)

        def batched():
            with gbamd.Batcher(eng, hi, lo, SEGMENT) as bt:
                for a, b in ranges:
                    bt.submit(a, b)
                bt.run_all()
                return bt.totals(), bt.batch_segments

        for _ in range(args.warmup):
            plain()
            batched()
        device_sync(lib)
        t_plain, t_batch, tot = [], [], []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            plain()
            device_sync(lib)
            t_plain.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            totals, kb = batched()
            device_sync(lib)
            t_batch.append(time.perf_counter() - t0)
            tot.append(totals)
        eng.close()
    k = len(t_batch)
    wall_b, wall_p = sum(t_batch) / k, sum(t_plain) / k
    kern, roof = harvest_roofline(tot)
    print(json.dumps({
        "metric": "batched vs plain per-range runs, G evens/s", "value": evens / wall_b / 1e9, "unit": "G evens/s",
        "n_gpus": 1, "steps": k, "warmup": args.warmup, "ms_per_step": 1e3 * wall_b, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: 256 seeded ranges (tests/golden/cfg4_ranges.txt)",
        "config": {"workload": "batched uneven ranges: 256 ranges, seed 1, starts log-uniform over [1e9, 5e18], "
                               "lengths log-uniform over [1e5, 4.675e8] integers",
                   "ranges": len(ranges), "evens": evens, "segment": SEGMENT, "batch_segments": kb,
                   "plain_value": evens / wall_p / 1e9, "plain_ms": 1e3 * wall_p, "batched_ms": 1e3 * wall_b,
                   "ratio": wall_p / wall_b, "kernel_ms": {n: 1e3 * v[0] / k for n, v in kern.items()},
                   "kernel_launches": {n: v[1] / k for n, v in kern.items()}, "bitmap_prepare_s": prep_s},
        "roofline": roof}), flush=True)


def bench_serving(args):
    import gbamd
    import numpy as np
    import serve

    reqs = serve.draw_requests(100, 5)
    evens = sum((b - a) // 2 + 1 for a, b in reqs)
    max_limit = serve.REQ_START_HI + 2 * serve.REQ_LEN_HI
    rates = (10, 30, 100, 300, 1000)
    out = []
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
            srv.close()
            lat = np.array([r["completed"] - r["arrival"] for r in rec])
            span = max(r["completed"] for r in rec) - min(arrivals)
            seq = np.array(serve.fifo_latencies(arrivals, service))
            seq_span = max(a + l for a, l in zip(arrivals, seq)) - min(arrivals)
            out.append({"rate": rate, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                        "p99_ms": 1e3 * float(np.percentile(lat, 99)), "throughput_gevens_s": evens / span / 1e9,
                        "plain_p50_ms": 1e3 * float(np.percentile(seq, 50)),
                        "plain_p99_ms": 1e3 * float(np.percentile(seq, 99)),
                        "plain_throughput_gevens_s": evens / seq_span / 1e9})
            print(json.dumps({"progress": "serving", **out[-1]}), file=sys.stderr, flush=True)
        eng.close()
    top = out[-1]
    print(json.dumps({
        "metric": "served p50/p99 latency (ms) and throughput", "value": top["throughput_gevens_s"],
        "unit": "G evens/s at the highest rate", "n_gpus": 1, "steps": len(rates), "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
        "data": "synthetic: 100 seeded requests (goldbach-l0_amd/serve.py draw_requests, seed 5)",
        "config": {"workload": "serving: Poisson arrivals at 10-1000 requests/s, start uniform in [1e12, 5e18], "
                               "length log-uniform in [1e4, 1e9] integers, 64 GB admission budget",
                   "requests": len(reqs), "evens": evens, "rates": out,
                   "plain_service_ms_mean": 1e3 * sum(service) / len(service), "bitmap_prepare_s": prep_s}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=("segment", "slice", "batched", "serving"), default="segment")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
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
    if args.workload != "segment":
        {"slice": bench_slice, "batched": bench_batched, "serving": bench_serving}[args.workload](args)
        return
    import gbamd

    lib = hip()
    with tempfile.TemporaryDirectory(dir=tmpfs()) as td, gbamd.env_scope(**stats_files(td)):
        eng = gbamd.Engine(LIMIT, SEGMENT)
        prep_s = eng.prepare(LIMIT, START)

        def step():
            fresh(td)
            return eng.run(LIMIT, SEGMENT, START)

        for _ in range(args.warmup):
            step()
        device_sync(lib)
        runs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            runs.append(step())          # synchronous: every segment collected and committed before it returns
        device_sync(lib)
        wall = time.perf_counter() - t0
        init_s = eng.init_seconds
        eng.close()

    k = len(runs)
    per_step = lambda key: sum(r[key] for r in runs) / k
    stage = {"large": per_step("large_seconds"), "sieve": per_step("sieve_seconds"),
             "check": per_step("verify_seconds") - per_step("count_seconds"), "count": per_step("count_seconds")}
    kern, roof = harvest_roofline(runs)
    dom = roof["kernel"]
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
                   "stage_ms": {n: 1e3 * v for n, v in stage.items()},
                   "kernel_ms": {n: 1e3 * v[0] / k for n, v in kern.items()},
                   "gpu_loop_ms": 1e3 * per_step("gpu_seconds"),
                   "host_plan_ms": 1e3 * per_step("host_plan_seconds"),
                   "engine_init_s": init_s, "bitmap_prepare_s": prep_s},
        "roofline": {**{k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac")},
                     "traffic": traffic() if dom == "k_harvest" else None,
                     **{k: roof[k] for k in ("kernel", "bytes_per_launch", "launch_ms")}},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
