import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG = ROOT / "goldbach-l0_amd"
sys.path.insert(0, str(PKG))
sys.path.insert(0, str(ROOT / "oracle"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture
def stats_env(tmp_path, monkeypatch):
    """Statistics files of a run isolated in a temp dir (fresh state)."""
    monkeypatch.setenv("GB_STATS_STATE", str(tmp_path / "st.state"))
    monkeypatch.setenv("GB_STATS_LOG", str(tmp_path / "st.log"))
    monkeypatch.setenv("GB_RECORDS_FILE", str(tmp_path / "records.txt"))
    monkeypatch.setenv("GB_STATS_LIVE", str(tmp_path / "live.json"))
    monkeypatch.setenv("GB_RIDGE_FILE", str(tmp_path / "ridges.csv"))
    monkeypatch.setenv("GB_RECORD_HOOK", "")
    for k in ("GB_STATS", "GB_GAPS", "GB_STATS_SAMPLE", "GB_WIN_KB", "GB_WIN_P", "GB_WIN_P_LONG", "GB_KA", "GB_BATCH",
              "GB_HL_MODE", "GB_OVERLAP", "GB_TUNE_FILE", "GB_HL_WHEEL_HITS", "GB_HL_PERSIST", "GB_HL_PERSIST_WG", "GB_HL_CHUNKS_CU", "GB_HL_HITQ",
              "GB_PHASEB_SLOTS", "GB_HL_REGION", "GB_GRAPH", "GB_CHECKB_SPLIT"):
        monkeypatch.delenv(k, raising=False)
    return tmp_path
