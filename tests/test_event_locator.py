"""The FP32 tracer's event locator (safeguarded Newton iteration on a step's dense output,
sirius_amd/csrc/geodesic.hpp) against bisection of the same polynomial, in a host build of the
header: tests/cpp/event_locator_check.cpp, compiled here with g++. Crossings of the disk plane and
of spheres of the Kerr radius; the bars are stated in that file."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "event_locator_check.cpp")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the HIP headers")
def test_newton_locator_matches_bisection(tmp_path):
    exe = str(tmp_path / "event_locator_check")
    b = subprocess.run(["g++", "-std=c++17", "-O2", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", SRC, "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout
