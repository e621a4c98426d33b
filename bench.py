"""Throughput benchmark of the MI355X verifier on one GPU.

Workload (one step): one full production segment range at 4e18, every even
n in [4e18, 4e18 + 467,499,998] (233,750,000 evens) through the whole
pipeline with exact statistics, as the reference runs that range (two
half segments by its at-least-two-segments rule). value = evens verified
per second of GPU-loop time (the reference's evens_per_sec convention:
total evens / segment-loop wall seconds, init excluded).

Prints one JSON line (see the fields below); --gpus above 1 is refused: this
build runs on one GPU (disjoint ranges per GPU would be independent runs).
"""
import argparse
import json
import os
import pathlib
import subprocess
import sys
import tempfile
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "goldbach-l0_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

START = 4_000_000_000_000_000_000
LIMIT = START + 467_499_998
SEGMENT = 467_500_000
HBM_PEAK_GBS = 8000.0


def cpu_baseline():
    """The oracle's per-even Miller-Rabin restatement of stats_ref (the
    reference's only CPU implementation of this path) on a bounded sample of
    the same range, all host cores (at most 16)."""
    import compare

    cores = min(16, os.cpu_count() or 1)
    n_evens = 100_000 * cores
    a = START
    b = START + 2 * (n_evens - 1)
    t0 = time.perf_counter()
    compare.oracle_ref(a, b, cores)
    dt = time.perf_counter() - t0
    return {"value": n_evens / dt / 1e9, "unit": "G evens/s", "cores": cores, "kind": "port",
            "sample": f"stats_ref restatement (oracle/gb_oracle.c ref) over [{a}, {b}]: {n_evens} evens, "
                      f"{dt:.1f} s wall on {cores} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; a multi-GPU campaign is disjoint ranges per GPU, run as "
              "independent processes (no collective is involved)", file=sys.stderr)
        sys.exit(2)
    import gbamd

    evens = (LIMIT - START) // 2 + 1
    with tempfile.TemporaryDirectory() as td, gbamd.env_scope(
            GB_STATS_STATE=f"{td}/st.state", GB_STATS_LOG=f"{td}/st.log", GB_RECORDS_FILE=f"{td}/rec.txt",
            GB_STATS_LIVE=f"{td}/live.json", GB_RECORD_HOOK=""):
        # the engine holds the base primes and tables in HBM: init is outside
        # the timed steps (the reference's convention excludes init too)
        eng = gbamd.Engine(LIMIT, SEGMENT)

        def step():
            for f in ("st.state", "st.log", "st.state.gaps"):
                pathlib.Path(td, f).unlink(missing_ok=True)
            return eng.run(LIMIT, SEGMENT, START)

        for _ in range(args.warmup):
            step()
        runs = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            runs.append(step())          # synchronous: every segment collected before it returns
        wall = time.perf_counter() - t0
        init_s = eng.init_seconds
        eng.close()
    gpu_s = wall
    stage = {k: sum(r[k] for r in runs) / len(runs) for k in ("sieve_seconds", "large_seconds", "verify_seconds",
                                                               "count_seconds")}
    segs = runs[-1]["segments"]
    # dominant stage and its algorithmic HBM bytes per launch
    dom = max(stage, key=stage.get)
    q_odds = (LIMIT - 3 - (START - 9781)) // 2 + 1
    map_bytes = q_odds * 8 / 15 / 8                      # class-map bits per run, as bytes
    nlarge = runs[-1]["total_large_items"] / segs
    per_run_bytes = {
        "sieve_seconds": map_bytes,                       # one store of every map word
        "large_seconds": 4 * nlarge * segs,               # the large-prime list, read once per segment
        "verify_seconds": map_bytes * 2,                  # checker + counter: one read each
        "count_seconds": map_bytes,
    }[dom]
    achieved = per_run_bytes / stage[dom] / 1e9
    out = {
        "metric": "G evens/s at 4e18 w/ exact stats",
        "value": evens * len(runs) / gpu_s / 1e9,
        "unit": "G evens/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * gpu_s / len(runs),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: the integer range itself (no dataset)",
        "config": {"workload": "one production segment range at 4e18: evens in [4e18, 4e18 + 467,499,998], "
                               "233,750,000 evens, 2 segments, exact statistics and gaps",
                   "start": START, "limit": LIMIT, "segment": SEGMENT, "segments": segs,
                   "stage_ms": {k.replace("_seconds", ""): 1e3 * v for k, v in stage.items()},
                   "gpu_loop_ms_per_step": 1e3 * sum(r["gpu_seconds"] for r in runs) / len(runs),
                   "engine_init_s": init_s},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel_stage": dom.replace("_seconds", "")},
        "cpu_baseline": None if args.no_cpu else cpu_baseline(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
