"""Tune files (the autotuner's output, read through the C ABI without a
GPU): band selection by the engine's limit, strict parsing, and the
environment's precedence over tuned defaults is the knob reader's (the GPU
tests run engines with a tune file)."""
import pytest

import gbamd


def _write(tmp_path, text):
    p = tmp_path / "tune.txt"
    p.write_text(text)
    return p


def test_band_selected_by_limit(tmp_path):
    p = _write(tmp_path, "# comment\n"
                         "band 1000000000000 10000000000000 GB_WIN_P=262144 GB_KA=64\n"
                         "band 1000000000000000000 10000000000000000000 GB_WIN_P=1048576 GB_HL_MODE=counted  # note\n")
    try:
        assert gbamd.tune_load(p, 4_000_000_000_000_000_000) == "GB_HL_MODE=counted GB_WIN_P=1048576"
        assert gbamd.tune_load(p, 5_000_000_000_000) == "GB_KA=64 GB_WIN_P=262144"
        assert gbamd.tune_load(p, 10**15) == ""
    finally:
        gbamd.tune_load(None, 0)


@pytest.mark.parametrize("text,msg", [
    ("band 1 10 GB_NOPE=3\n", "unknown knob"),
    ("band 10 1 GB_KA=64\n", "empty band"),
    ("bend 1 10\n", "expected 'band'"),
    ("band 1 x GB_KA=64\n", "not a number"),
    ("band 1 10 GB_KA\n", "not KNOB=value"),
])
def test_malformed_files_are_refused(tmp_path, text, msg):
    p = _write(tmp_path, text)
    with pytest.raises(gbamd.VerificationError, match=msg):
        gbamd.tune_load(p, 5)


def test_missing_file_is_refused(tmp_path):
    with pytest.raises(gbamd.VerificationError, match="cannot read"):
        gbamd.tune_load(tmp_path / "none.txt", 5)
