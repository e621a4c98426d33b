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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
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
    import gbamd

    lib = hip()
    # the run's statistics files are thrown away after each step: on tmpfs
    # their fsyncs measure memory, not whatever disk the host has
    shm = "/dev/shm" if os.path.isdir("/dev/shm") and os.access("/dev/shm", os.W_OK) else None
    with tempfile.TemporaryDirectory(dir=shm) as td, gbamd.env_scope(
            GB_STATS_STATE=f"{td}/st.state", GB_STATS_LOG=f"{td}/st.log", GB_RECORDS_FILE=f"{td}/rec.txt",
            GB_STATS_LIVE=f"{td}/live.json", GB_RIDGE_FILE=f"{td}/ridges.csv", GB_RECORD_HOOK=""):
        eng = gbamd.Engine(LIMIT, SEGMENT)
        prep_s = eng.prepare(LIMIT, START)

        def step():
            for f in ("st.state", "st.log", "st.state.gaps"):
                pathlib.Path(td, f).unlink(missing_ok=True)
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
    kern = {"k_harvest": (sum(r["harvest_kernel_seconds"] for r in runs), sum(r["harvest_launches"] for r in runs),
                          sum(r["harvest_bytes"] for r in runs)),
            "k_window": (sum(r["window_kernel_seconds"] for r in runs), sum(r["window_launches"] for r in runs),
                         sum(r["window_bytes"] for r in runs))}
    dom = max(kern, key=lambda n: kern[n][0])
    sec, launches, nbytes = kern[dom]
    launch_s = sec / max(launches, 1)
    bytes_per_launch = nbytes / max(launches, 1)
    achieved = bytes_per_launch / launch_s / 1e9 if launch_s > 0 else 0.0
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic() if dom == "k_harvest" else None,
                     "kernel": dom, "bytes_per_launch": bytes_per_launch, "launch_ms": 1e3 * launch_s},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
