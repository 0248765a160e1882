"""The log10 of the per-turn MEGNO sums (henon_map_cpp_amd/csrc/hm_log10.h, the
fdlibm method) built for the host with g++ and checked against the C
library's log10: at most 2 units in the last place over the normal range and
around 1, correctly rounded for most arguments, and the library's results for
zero, negative, infinite and NaN arguments. The GPU tests hold the MEGNO sums
that use it to the oracle's (which uses the C library's log10)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "henon_map_cpp_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_log10_fdl_against_c_library(tmp_path):
    exe = str(tmp_path / "log10_check")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-I", CSRC,
                    os.path.join(HERE, "log10_check.cpp"), "-o", exe], check=True)
    worst, exact, ok = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert float(worst) <= 2.0
    assert float(exact) > 0.85
    assert ok == "1"
