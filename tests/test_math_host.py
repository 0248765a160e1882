"""The device math of the hot loops, built for the host with g++ and checked
against the C library: log10 of the per-turn MEGNO sums
(henon_map_cpp_amd/csrc/hm_log10.h), by the fdlibm method and branch-free,
stays within 2 units in the last place of the library's result, is correctly
rounded for most arguments, and gives the library's results for zero,
negative, infinite and NaN arguments. The GPU tests hold the MEGNO sums that
use it to the oracle's (which uses the C library)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "henon_map_cpp_amd", "csrc")


def run_check(tmp_path, name):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-I", CSRC, os.path.join(HERE, name + ".cpp"),
                    "-o", exe], check=True)
    return subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()[-3:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_log10_fdl_against_c_library(tmp_path):
    worst, exact, ok = run_check(tmp_path, "log10_check")
    assert float(worst) <= 2.0
    assert float(exact) > 0.85
    assert ok == "1"
