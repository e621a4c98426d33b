"""CPU-side checks of the product: the C-ABI library loads and exports every
entry point its headers declare, the host API the reference's command-line
driver calls (src/main.cpp of the reference) compiles and links against
this build's headers and library (tests/cli_api_check.cpp), the CLI keeps
the reference's usage contract (tests/cli_tests.sh: exit 2 on usage errors),
and a run without a GPU fails closed."""
import ctypes
import pathlib
import re
import subprocess

import pytest

import gbamd
from conftest import ROOT


def _declared(header):
    text = (ROOT / "include" / header).read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(gb[ah]_\w+)\s*\(", text, re.M)))


@pytest.mark.parametrize("header", ["gbh.h", "goldbach_amd.h"])
def test_library_exports_every_declared_symbol(header):
    names = _declared(header)
    assert len(names) >= 3
    lib = ctypes.CDLL(str(gbamd.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_library_loads_and_reports_version():
    assert gbamd.version().startswith("2.8.0")


def test_reference_cli_host_api_links(tmp_path):
    pkg = ROOT / "goldbach-l0_amd"
    exe = tmp_path / "cli_api_check"
    r = subprocess.run(["g++", "-std=c++23", "-O1", "-Wall", "-Wextra", f"-I{ROOT / 'include'}",
                        f"-I{pkg / 'csrc' / 'host'}", "-o", str(exe), str(ROOT / "tests" / "cli_api_check.cpp"),
                        f"-L{pkg}", "-lgoldbach_amd", f"-Wl,-rpath,{pkg}", "-pthread"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "0 failed" in r.stdout, r.stdout + r.stderr


def _cli(*args):
    return subprocess.run([str(gbamd.CLI_PATH), *args], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("args", [
    ["--limit", "1e9x"], ["--limit", "010e"], ["--start", "10", "--limit", "8"], ["--cpu-check", "0"],
    ["--bogus"], ["--split", "5000000000", "--limit", "100"], ["--limit"],
])
def test_cli_usage_errors_exit_2(args):
    r = _cli(*args)
    assert r.returncode == 2, (args, r.stdout, r.stderr)


def test_cli_cpu_check_and_number_syntax():
    r = _cli("--cpu-check", "2e5")
    assert r.returncode == 0 and "CPU check up to 200000: PASS" in r.stdout


def test_cli_removed_knob_refused(monkeypatch):
    monkeypatch.setenv("GB_HITLIST", "1")
    r = _cli("--start", "1e6", "--limit", "2e6")
    assert r.returncode == 1 and "removed knob" in r.stderr


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-GPU failure path")
def test_run_without_gpu_fails_closed(stats_env):
    with pytest.raises(gbamd.VerificationError):
        gbamd.run_gpu(2_000_000, 467_500_000, start_even=1_000_000)
    assert not (stats_env / "st.state").exists()
