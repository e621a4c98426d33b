"""Parity checks of a verifier run against the CPU oracle (TEST INFRASTRUCTURE:
used by tests/, the smoke check and bench.py's CPU-baseline leg only).

``compare_ref`` decides whether a run's GBSTATS log line plus its state
file describe the same window as a stats_ref output line (the reference's
criteria, tests/stats_compare.py:21-88 of the reference): both sides are
reduced to one summary of comparable facts, and the facts the run can state
exactly (the full p(n) histogram only when every segment was sampled, the
walked gap histogram and moments only when every segment was walked) are
compared.

``oracle_gbstats`` runs this repository's C oracle (gb_oracle.c), which
writes the GBSTATS statistics fields of a fresh-state run from the
definitions, so a run's line can be compared byte for byte.
"""
from __future__ import annotations

import os
import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
ORACLE = HERE / "gb_oracle"


def kv(line: str) -> dict:
    """The key=value words of a line."""
    out = {}
    for word in line.split():
        key, eq, val = word.partition("=")
        if eq:
            out[key] = val
    return out


def odd_primes(n: int = 9781) -> list:
    """The odd primes up to n."""
    flags = bytearray([1]) * (n + 1)
    flags[0:2] = b"\x00\x00"
    for p in range(2, int(n ** 0.5) + 1):
        if flags[p]:
            flags[p * p::p] = bytes(len(range(p * p, n + 1, p)))
    return [p for p in range(3, n + 1) if flags[p]]


def _counts(text) -> list:
    """A comma list of counts with empty entries for zero ("" or "-": none)."""
    if text in (None, "", "-"):
        return []
    return [int(x) if x else 0 for x in text.split(",")]


def _pairs(text, sep=";") -> dict:
    """'a:b;c:d' as {a: b, c: d} ("-" or "": empty)."""
    if text in (None, "", "-"):
        return {}
    return {int(a): int(b) for a, b in (item.split(":") for item in text.split(sep))}


def _padded(v: list, n: int) -> list:
    return list(v[:n]) + [0] * (n - len(v))


def _ref_summary(ref_line: str, primes: list, tail_t: int) -> dict:
    """The facts of a stats_ref output line."""
    r = kv(ref_line)
    hist = _padded(_counts(r["hist"]), len(primes))
    mom = r.get("gmom", "").split(",")
    return {
        "totals": {k: r.get(k) for k in ("evens", "primes", "twins", "primes30", "twins30")},
        "tail": {p: c for p, c in zip(primes, hist) if c and p >= tail_t},
        "first": {p: n for p, n in _pairs(r["first"]).items() if p >= tail_t},
        "hist": hist,
        "gaps": {"gaps": r.get("gaps"), "gd1": mom[0], "gtail": r.get("gtail")},
        "gmax": r.get("gmax"),
        "walk_hist": _counts(r.get("ghist")),
        "walk_moments": (r.get("gaps"), r.get("gmom")),
    }


def _run_summary(gbstats_line: str, state_text: str, primes: list) -> dict:
    """The facts of a run's GBSTATS line and its (fresh-campaign) state file."""
    g = kv(gbstats_line)
    tail = {}
    if g["tail"] != "-":
        p0, _, lst = g["tail"].partition(":")
        at = primes.index(int(p0))
        tail = {primes[at + i]: c for i, c in enumerate(_counts(lst)) if c}
    sampled = [0] * len(primes)
    for line in state_text.splitlines():
        if line.startswith("sampled "):
            for i, c in enumerate(_counts(line.split(" ", 2)[2])):
                sampled[i] += c
    walked_evens, _, walked_gaps = g.get("gwalk", "0:0").partition(":")
    return {
        "totals": {k: g.get(k) for k in ("evens", "primes", "twins", "primes30", "twins30")},
        "tail": tail,
        "first": _pairs(g["new_first"]),
        "hist": sampled,
        "hist_total": int(g["evens_sampled"]),
        "all_sampled": g["evens"] == g["evens_sampled"],
        "has_gaps": "gaps" in g,
        "gaps": {"gaps": g.get("gaps"), "gd1": g.get("gd1"), "gtail": g.get("gtail")},
        "gmax": g.get("gmax"),
        "all_walked": walked_evens == g["evens"],
        "walk_hist": _counts(g.get("gsamp")),
        "walk_moments": (walked_gaps, g.get("gsmom")),
    }


def compare_ref(gbstats_line: str, state_text: str, ref_line: str) -> list:
    """Mismatches between a run (its GBSTATS line + state file, fresh state)
    and the stats_ref output of the same range; [] when they agree."""
    primes = odd_primes()
    tail_t = int(kv(gbstats_line)["tail_t"])
    run = _run_summary(gbstats_line, state_text, primes)
    ref = _ref_summary(ref_line, primes, tail_t)
    bad = []

    def differ(what, a, b):
        if a != b:
            if isinstance(a, dict) and isinstance(b, dict):
                delta = sorted(set(a.items()) ^ set(b.items()))[:6]
                bad.append(f"{what} differ: {delta}")
            elif isinstance(a, list) and isinstance(b, list):
                n = max(len(a), len(b))
                a2, b2 = _padded(a, n), _padded(b, n)
                bad.append(f"{what} differ at " + str([(i, x, y) for i, (x, y) in enumerate(zip(a2, b2)) if x != y][:6]))
            else:
                bad.append(f"{what}: gpu {a} cpu {b}")

    for k in ref["totals"]:
        differ(k, run["totals"][k], ref["totals"][k])
    differ("tail counts", run["tail"], ref["tail"])
    differ("first occurrences", run["first"], ref["first"])
    if sum(run["hist"]) != run["hist_total"]:
        bad.append(f"sampled histogram total {sum(run['hist'])} != evens_sampled {run['hist_total']}")
    if run["all_sampled"]:
        differ("full histogram", run["hist"], ref["hist"][:len(run["hist"])])
    if run["has_gaps"]:
        for k in ref["gaps"]:
            differ(k, run["gaps"][k], ref["gaps"][k])
        if run["gmax"] != "-":
            differ("gmax", run["gmax"], ref["gmax"])
        if run["all_walked"]:
            differ("walked gap histogram", run["walk_hist"], ref["walk_hist"])
            differ("walked gap moments", run["walk_moments"], ref["walk_moments"])
            differ("gmax (every segment walked)", run["gmax"], ref["gmax"])
    return bad


def load_fixture(path) -> tuple:
    """(start, limit, stats_ref line) of a golden fixture file."""
    start, limit, line = pathlib.Path(path).read_text().strip().split(" ", 2)
    return int(start), int(limit), line


def _threads(threads):
    return threads or min(16, os.cpu_count() or 1)


def oracle_gbstats(start: int, limit: int, segment: int, sample: int = 256, gaps: str = "exact",
                   threads: int | None = None) -> tuple:
    """(GBSTATS fields after "version=...", state histogram records) of a
    fresh-state run, from the C oracle."""
    out = subprocess.run([str(ORACLE), "gbstats", str(start), str(limit), str(segment), str(sample), gaps,
                          str(_threads(threads))], check=True, capture_output=True, text=True).stdout
    first, *rest = out.splitlines()
    return first, rest


def oracle_ref(start: int, limit: int, threads: int | None = None) -> str:
    """The oracle's stats_ref-format line of [start, limit]."""
    return subprocess.run([str(ORACLE), "ref", str(start), str(limit), str(_threads(threads))], check=True,
                          capture_output=True, text=True).stdout.strip()


def strip_version(gbstats_line: str) -> str:
    """The GBSTATS line from its "start=" field on."""
    return "start=" + gbstats_line.strip().partition(" start=")[2]


def state_hist_records(state_text: str) -> list:
    """The state file's histogram records."""
    return [line for line in state_text.splitlines() if line.startswith(("sampled ", "tailcls ", "tailagg "))]
