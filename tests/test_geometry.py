"""Host-only check of the stage-1 tile schedule of shifted stripes (tests/geometry/): the span
helpers of projekt-2012_amd/csrc/sw_host.h against a brute-force scan, every lane's columns
covered exactly once, no two tiles of a launch touching one hand-off location when one of them
writes it, and every read finding the value the right tile wrote in an earlier launch.  Built
with g++ against the HIP headers (no device code, no device needed)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "geometry", "shifted_check.cpp")
HIP_INCLUDE = "/opt/rocm/include"


def test_shifted_stripe_schedule(tmp_path):
    if shutil.which("g++") is None or not os.path.isdir(HIP_INCLUDE):
        pytest.skip("g++ or the HIP headers are not available")
    exe = str(tmp_path / "shifted_check")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + HIP_INCLUDE,
                        "-I" + os.path.join(ROOT, "projekt-2012_amd", "csrc"), "-o", exe, SRC],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout[-2000:]
