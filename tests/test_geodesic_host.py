"""Host builds of the tracer's header (sirius_amd/csrc/geodesic.hpp), compiled here with g++:
  - event_locator_check.cpp: the FP32 tracer's event locator (safeguarded Newton iteration on a
    step's dense output) against bisection of the same polynomial, on crossings of the disk
    plane and of spheres of the Kerr radius;
  - warp_shape_check.cpp: the Alcubierre shape function and its derivative by one exponential
    against the two-tanh definition.
The bars are stated in each program."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the HIP headers")
@pytest.mark.parametrize("program", ["event_locator_check.cpp", "warp_shape_check.cpp"])
def test_host_build_of_the_tracer_header(tmp_path, program):
    exe = str(tmp_path / program.replace(".cpp", ""))
    b = subprocess.run(["g++", "-std=c++17", "-O2", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                        os.path.join(HERE, "cpp", program), "-o", exe], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout
