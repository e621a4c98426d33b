"""Autotuner of the pipeline's shape knobs.

Sweeps, one knob at a time from the current best (coordinate descent, two
passes), the choices that shape the pipeline on a sample workload:

    GB_WIN_KB   LDS window of the sieve (0: sized per run; 32, 64, 128 KB)
    GB_WIN_P    bound between the window primes and the large primes
    GB_WIN_P_LONG  long runs' bound: their large wheel primes below it are
                sieved in the windows too
    GB_KA       phase-A candidates per even class
    GB_BATCH    segments per batch (ranges of many segments)
    GB_HL_MODE  harvest form of long batches (direct or counted)
    GB_OVERLAP  sieve phase beside the harvest on a second stream
    GB_HL_WHEEL_HITS  wheel walk instead of random bitmap reads below this
                many hits per prime
    GB_GRAPH    each batch's launches as one captured HIP graph, or issued
                one by one

Each candidate runs the sample range on a fresh engine (one warmup, then
the best of `--reps` timed runs; evens per second of wall time). Every choice
is exact (the statistics do not depend on it), so the tuner only weighs
speed. The best setting is written as the band of the sample's magnitude
(one decade: [10^k, 10^(k+1))) into the tune file, replacing an earlier line
for that band; an engine then takes the band of its limit from
GB_TUNE_FILE for every knob the environment does not set.

The sample should hold several batches (the default, 160 full segments, is
ten of 16): a sample of one batch cannot show the overlap of consecutive
batches and favours smaller batches (16 segments: GB_BATCH=8 first; 160:
GB_BATCH=16 by 7%).
    python goldbach-l0_amd/autotune.py --start 4.5e18 --out tune.txt
"""
from __future__ import annotations

import argparse
import math
import os
import pathlib
import sys
import tempfile
import time

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent))
import gbamd  # noqa: E402

SPACE = {
    "GB_WIN_KB": ["0", "32", "64", "128"],
    "GB_WIN_P": ["262144", "524288", "1048576"],
    "GB_WIN_P_LONG": ["524288", "1048576", "2097152"],
    "GB_KA": ["64", "96", "128"],
    "GB_BATCH": ["8", "16"],
    "GB_HL_MODE": ["auto", "counted"],
    "GB_OVERLAP": ["1", "0"],
    "GB_HL_WHEEL_HITS": ["0", "2", "4", "8"],
    "GB_GRAPH": ["0", "1"],
}
DEFAULT = {"GB_WIN_KB": "0", "GB_WIN_P": "524288", "GB_WIN_P_LONG": "1048576", "GB_KA": "96", "GB_BATCH": "16",
           "GB_HL_MODE": "auto", "GB_OVERLAP": "1", "GB_HL_WHEEL_HITS": "2", "GB_GRAPH": "0"}


def band_of(limit: int):
    k = int(math.floor(math.log10(limit)))
    return 10**k, 10**(k + 1)


def measure(start: int, limit: int, segment: int, knobs: dict, reps: int, td: str) -> float:
    """Evens per second of the best of `reps` runs with `knobs`."""
    with gbamd.env_scope(**knobs):
        eng = gbamd.Engine(limit, segment)
        try:
            eng.prepare(limit, start)
            best = math.inf
            for k in range(reps + 1):
                for f in ("st.state", "st.log", "st.state.gaps"):
                    pathlib.Path(td, f).unlink(missing_ok=True)
                t = time.perf_counter()
                eng.run(limit, segment, start)
                w = time.perf_counter() - t
                if k:
                    best = min(best, w)
        finally:
            eng.close()
    return ((limit - start) // 2 + 1) / best


def write_band(path: pathlib.Path, lo: int, hi: int, knobs: dict, note: str):
    lines = path.read_text().splitlines() if path.exists() else []
    keep = [l for l in lines if not l.startswith(f"band {lo} {hi} ")]
    if not any(l.startswith("#") for l in keep):
        keep.insert(0, "# goldbach-l0_amd tune file: band <lo> <hi> KNOB=value ... (autotune.py)")
    keep.append(f"band {lo} {hi} " + " ".join(f"{k}={v}" for k, v in sorted(knobs.items())) + f"  # {note}")
    path.write_text("\n".join(keep) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--start", type=float, required=True, help="first even of the sample range")
    ap.add_argument("--length", type=float, default=7.48e10, help="integers in the sample range")
    ap.add_argument("--segment", type=int, default=467_500_000)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--out", default="goldbach_tune.txt")
    args = ap.parse_args(argv)
    start = int(args.start) & ~1
    limit = (start + int(args.length) - 1) & ~1
    with tempfile.TemporaryDirectory() as td, gbamd.env_scope(
            GB_STATS_STATE=f"{td}/st.state", GB_STATS_LOG=f"{td}/st.log", GB_RECORDS_FILE=f"{td}/rec.txt",
            GB_STATS_LIVE=f"{td}/live.json", GB_RIDGE_FILE=f"{td}/ridges.csv", GB_RECORD_HOOK="", GB_TUNE_FILE=None):
        best = dict(DEFAULT)
        best_v = measure(start, limit, args.segment, best, args.reps, td)
        print(f"defaults: {best_v / 1e9:.2f} G evens/s", flush=True)
        for p in range(args.passes):
            for knob, choices in SPACE.items():
                for c in choices:
                    if c == best[knob]:
                        continue
                    trial = dict(best, **{knob: c})
                    try:
                        v = measure(start, limit, args.segment, trial, args.reps, td)
                    except gbamd.VerificationError as e:      # a shape this range cannot take
                        print(f"  {knob}={c}: refused ({e})", flush=True)
                        continue
                    print(f"  {knob}={c}: {v / 1e9:.2f} G evens/s", flush=True)
                    if v > best_v * 1.01:                     # a clear gain only
                        best, best_v = trial, v
            print(f"pass {p + 1}: {best_v / 1e9:.2f} G evens/s with {best}", flush=True)
    lo, hi = band_of(limit)
    write_band(pathlib.Path(args.out), lo, hi, best,
               f"{best_v / 1e9:.2f} G evens/s on [{start}, {limit}], segment {args.segment}")
    print(f"wrote band [{lo}, {hi}) to {args.out}", flush=True)


if __name__ == "__main__":
    main()
