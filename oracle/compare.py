"""Parity checks of a verifier run against the CPU oracle (TEST INFRASTRUCTURE:
used by tests/, the smoke check and bench.py's CPU-baseline leg only).

``compare_ref`` restates the reference's field comparison between a GBSTATS
log line + state file and the stats_ref output line (tests/stats_compare.py:
21-88): exact evens / primes / twins by residue, the p(n) tail and its first
occurrences, the sampled-histogram total, the full histogram when every
segment was sampled, and the gap count / sum / tail / maximum (full gap
histogram and moments when every segment was walked).

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
REF = HERE / "_ref" / "stats_ref"


def kv(line: str) -> dict:
    return dict(t.split("=", 1) for t in line.split() if "=" in t)


def odd_primes(n: int = 9781) -> list:
    comp = bytearray(n + 1)
    out = []
    for p in range(3, n + 1, 2):
        if not comp[p]:
            out.append(p)
            for m in range(p * p, n + 1, 2 * p):
                comp[m] = 1
    return out


def dense(v: str) -> list:
    if v in (None, "", "-"):
        return []
    return [int(x) if x else 0 for x in v.split(",")]


def compare_ref(gbstats_line: str, state_text: str, ref_line: str) -> list:
    """Mismatches between a run (its GBSTATS line + state file, fresh state)
    and the stats_ref output of the same range; [] when they agree."""
    g = kv(gbstats_line)
    c = kv(ref_line)
    pr = odd_primes()
    T = int(g["tail_t"])
    bad = []
    for k in ("evens", "primes", "twins", "primes30", "twins30"):
        if g.get(k) != c.get(k):
            bad.append(f"{k}: gpu {g.get(k)} cpu {c.get(k)}")
    ch = dense(c["hist"])
    ch += [0] * (len(pr) - len(ch))
    cpu_tail = {pr[i]: v for i, v in enumerate(ch) if v and pr[i] >= T}
    gt = {}
    if g["tail"] != "-":
        p0, lst = g["tail"].split(":")
        i0 = pr.index(int(p0))
        for i, x in enumerate(lst.split(",")):
            if x:
                gt[pr[i0 + i]] = int(x)
    if gt != cpu_tail:
        bad.append(f"tail differs: {sorted(set(gt.items()) ^ set(cpu_tail.items()))[:6]}")
    acc = {}
    for line in state_text.splitlines():
        if line.startswith("sampled "):
            _, _, lst = line.split(" ", 2)
            for i, x in enumerate(lst.split(",") if lst else []):
                if x:
                    acc[i] = acc.get(i, 0) + int(x)
    gh = [acc.get(i, 0) for i in range(len(pr))]
    if sum(gh) != int(g["evens_sampled"]):
        bad.append(f"sampled histogram total {sum(gh)} != evens_sampled {g['evens_sampled']}")
    if g["evens"] == g["evens_sampled"] and gh != ch[:len(gh)]:
        bad.append("full histogram differs: " + str([(pr[i], a, b) for i, (a, b) in enumerate(zip(gh, ch)) if a != b][:6]))
    gf = {} if g["new_first"] == "-" else {int(a): int(b) for a, b in (x.split(":") for x in g["new_first"].split(";"))}
    cf = {int(a): int(b) for a, b in (x.split(":") for x in c["first"].split(";"))}
    if gf != {p: n for p, n in cf.items() if p >= T}:
        bad.append(f"first occurrences differ: {sorted(set(gf.items()) ^ set((p, n) for p, n in cf.items() if p >= T))[:6]}")
    if "gaps" in g:
        cm = c.get("gmom", "").split(",")
        for name, gv, cv in (("gaps", g.get("gaps"), c.get("gaps")), ("gd1", g.get("gd1"), cm[0]),
                             ("gtail", g.get("gtail"), c.get("gtail"))):
            if gv != cv:
                bad.append(f"{name}: gpu {gv} cpu {cv}")
        if g.get("gmax") != "-" and g.get("gmax") != c.get("gmax"):
            bad.append(f"gmax: gpu {g.get('gmax')} cpu {c.get('gmax')}")
        we, wn = g.get("gwalk", "0:0").split(":")
        if we == g["evens"]:
            gs, gh2 = dense(g.get("gsamp")), dense(c.get("ghist"))
            n = max(len(gs), len(gh2))
            gs, gh2 = gs + [0] * (n - len(gs)), gh2 + [0] * (n - len(gh2))
            if gs != gh2:
                bad.append("full gap histogram differs: " + str([(2 * i, a, b) for i, (a, b) in enumerate(zip(gs, gh2)) if a != b][:6]))
            if g.get("gsmom") != c.get("gmom") or wn != c.get("gaps"):
                bad.append(f"walked moments: gpu {wn} {g.get('gsmom')} cpu {c.get('gaps')} {c.get('gmom')}")
            if g.get("gmax") != c.get("gmax"):
                bad.append(f"gmax with every segment walked: gpu {g.get('gmax')} cpu {c.get('gmax')}")
    return bad


def load_fixture(path) -> tuple:
    """(start, limit, stats_ref line) of a golden fixture file."""
    a, b, rest = pathlib.Path(path).read_text().strip().split(" ", 2)
    return int(a), int(b), rest


def oracle_gbstats(start: int, limit: int, segment: int, sample: int = 256, gaps: str = "exact",
                   threads: int | None = None) -> tuple:
    """(GBSTATS fields after "version=...", state histogram records) of a
    fresh-state run, from the C oracle."""
    nt = threads or min(16, os.cpu_count() or 1)
    out = subprocess.run([str(ORACLE), "gbstats", str(start), str(limit), str(segment), str(sample), gaps, str(nt)],
                         check=True, capture_output=True, text=True).stdout
    lines = out.splitlines()
    return lines[0], lines[1:]


def oracle_ref(start: int, limit: int, threads: int | None = None) -> str:
    nt = threads or min(16, os.cpu_count() or 1)
    return subprocess.run([str(ORACLE), "ref", str(start), str(limit), str(nt)], check=True, capture_output=True,
                          text=True).stdout.strip()


def strip_version(gbstats_line: str) -> str:
    """The GBSTATS line after its "version=..." field."""
    head, _, rest = gbstats_line.strip().partition(" start=")
    return "start=" + rest


def state_hist_records(state_text: str) -> list:
    return [l for l in state_text.splitlines() if l.startswith(("sampled ", "tailcls ", "tailagg "))]
