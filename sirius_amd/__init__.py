"""MI355X-native geodesic-tracing backend for Sirius (HIP kernels behind a C-ABI).

The compute lives in libsirius_hip.so (sirius_amd/csrc, built for gfx950); this package is the
Python mirror of its IAccelerator/C-ABI surface used by the tests and the benchmark.
"""
from . import abi  # noqa: F401
from .accelerator import HipAccelerator, SiriusHipError, default_opts, lib  # noqa: F401
from .build import build  # noqa: F401

__all__ = ["abi", "HipAccelerator", "SiriusHipError", "default_opts", "lib", "build"]
