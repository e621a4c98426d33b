"""The host layer's C++ unit tests (tests/host_unit.cpp), built plain and with
AddressSanitizer + UBSan (host code only), each run to a clean exit: the
text codecs and file helpers, Miller-Rabin and the base-prime sieve, the
GBP2 prime cache, the ridge board and store, and the statistics' campaign
files (state v2, gap state v1, log rows, overlap refusal, damaged files)."""
import subprocess

import pytest

from conftest import PKG


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-j4", "-C", str(PKG), "host_unit"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return PKG / "build"


@pytest.mark.parametrize("binary", ["host_unit", "host_unit_san"])
def test_host_unit(built, binary):
    r = subprocess.run([str(built / binary)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 failed" in r.stdout
