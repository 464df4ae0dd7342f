"""The C++ IAccelerator implementation (HIPAccelerator, include/sirius_hip_accelerator.hpp) driven
by a compiled program as a Sirius tree would drive it (tests/cpp/hip_accelerator_check.cpp, built
by __graft_entry__.build()): launch from a LaunchConfig, the star field of
LaunchConfig::enableStarfield with ray bundles, and the image writers."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(__file__), "cpp", "hip_accelerator_check")


def test_cpp_wrapper_end_to_end(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: run __graft_entry__.build()")
    print("[case] C++ HIPAccelerator program", flush=True)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
