"""Engines with a tune file (GB_TUNE_FILE): the band's knobs shape the
pipeline (here the window-prime bound, visible as the run's split), the
environment overrides them, and the statistics are the untuned run's."""
import pytest

import compare
import gbamd

pytestmark = pytest.mark.gpu

A, B = 4_000_000_000_000_000_000, 4_000_000_000_002_000_000


def _run(stats_env):
    for f in ("st.state", "st.log", "st.state.gaps"):
        (stats_env / f).unlink(missing_ok=True)
    return gbamd.run_gpu(B, 700_000, start_even=A)


def test_tune_file_shapes_runs_without_changing_results(stats_env, monkeypatch, tmp_path):
    base = _run(stats_env)
    tune = tmp_path / "tune.txt"
    tune.write_text("band 1000000000000000000 10000000000000000000 GB_WIN_P=262144 GB_HL_MODE=counted\n")
    monkeypatch.setenv("GB_TUNE_FILE", str(tune))
    tuned = _run(stats_env)
    assert tuned["split_used"] < base["split_used"]            # fewer window primes: the file's bound
    assert compare.strip_version(tuned["stats_line"]) == compare.strip_version(base["stats_line"])
    monkeypatch.setenv("GB_WIN_P", "524288")                    # the environment wins
    env = _run(stats_env)
    assert env["split_used"] == base["split_used"]
    assert compare.strip_version(env["stats_line"]) == compare.strip_version(base["stats_line"])


@pytest.mark.parametrize("name,segment", [("w1e12", 1_000_000), ("w4e18", 2_000_000), ("w4e18l", 20_000_000)])
def test_shipped_tune_bands_match_stats_ref(stats_env, monkeypatch, name, segment):
    """The autotuned bands shipped with the package (tune_mi355x.txt) keep
    every statistic equal to the reference's fixture on windows of their
    decades."""
    from conftest import GOLDEN, PKG

    a, b, ref = compare.load_fixture(GOLDEN / f"ref_{name}.txt")
    monkeypatch.setenv("GB_TUNE_FILE", str(PKG / "tune_mi355x.txt"))
    print(f"tuned case {name} [{a}, {b}] segment {segment}", flush=True)
    r = gbamd.run_gpu(b, segment, start_even=a)
    assert r["leftovers"] == 0
    bad = compare.compare_ref(r["stats_line"], (stats_env / "st.state").read_text(), ref)
    assert not bad, bad
