"""Build recipe: compiles libsirius_hip.so for gfx950 (hipcc via sirius_amd/csrc/Makefile)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs=8, force=False):
    """Compile every HIP/C++ source into sirius_amd/lib/libsirius_hip.so. Always runs make
    (incremental); force=True rebuilds from scratch."""
    if force:
        subprocess.check_call(["make", "-C", CSRC, "-s", "clean"])
    subprocess.check_call(["make", "-C", CSRC, "-s", f"-j{jobs}"])
    return os.path.join(_HERE, "lib", "libsirius_hip.so")
