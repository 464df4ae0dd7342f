"""sirius_hip_trace_batch_noalloc (the trace with its scratch in a fixed device array) against
sirius_hip_trace_batch on the same inputs, through the C-ABI: the same kernels, so the same
results bit for bit. Two calls in a row, the second with a smaller batch, so the second runs on
the scratch the first left behind."""
import ctypes as C

import numpy as np
import pytest

from devmem import DeviceArray
from sirius_amd import abi
from sirius_amd.accelerator import HipAccelerator, default_opts

pytestmark = pytest.mark.gpu

# (field of SiriusHipRayOutput, dtype, components per ray)
FIELDS = (("cls", np.uint8, 1), ("flags", np.uint8, 1), ("theta", np.float32, 1), ("phi", np.float32, 1),
          ("lambda_", np.float32, 1), ("disk_r", np.float32, 1), ("redshift", np.float32, 1),
          ("null_drift", np.float32, 1), ("steps", np.uint32, 1), ("intensity", np.float32, -1),
          ("state", np.float64, 8))


def _device_out(n, n_bins, with_state):
    arrs = {}
    for name, dt, k in FIELDS:
        if name == "state" and not with_state:
            continue
        shape = (max(n_bins, 1), n) if k < 0 else ((n, k) if k > 1 else (n,))
        arrs[name] = DeviceArray(shape=shape, dtype=dt)
    ro = abi.RayOutput(on_device=1, **{k: v.data_ptr() for k, v in arrs.items()})
    return ro, arrs


def _trace(acc, entry, cams, metrics, opts, with_state):
    n = sum(c.width * c.height for c in cams)
    ro, arrs = _device_out(n, opts.n_bins, with_state)
    ca = (abi.Camera * len(cams))(*cams)
    ma = (abi.Metric * len(metrics))(*metrics)
    st = abi.Stats()
    fn = getattr(acc._L, entry)
    rc = fn(acc._h, ca, len(cams), ma, len(metrics), C.byref(opts), C.byref(ro), C.byref(st))
    assert rc == 0, (entry, rc, acc.getLastError())
    return {k: v.get() for k, v in arrs.items()}, st.as_dict()


def _same(a, b, n_bins):
    for name, _, k in FIELDS:
        if name not in a:
            continue
        x, y = a[name], b[name]
        if k < 0:
            x, y = x[:n_bins], y[:n_bins]
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name


@pytest.fixture(scope="module")
def own_acc():
    a = HipAccelerator()
    yield a
    a.close()


def test_noalloc_trace_equals_trace_batch_twice_second_smaller(own_acc):
    acc = own_acc
    # 1: the 1920x1080 Kerr a = 0.9 frame, FP32 path (gate, critical flags, persistent queue)
    cam = abi.camera(50.0, 75.0, 20.0, 1920, 1080, disk=(0.0, 30.0))
    m = abi.metric_ks(1.0, 0.9)
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    print("[case] noalloc 1920x1080 kerr a=0.9 fp32", flush=True)
    ref, st_ref = _trace(acc, "sirius_hip_trace_batch", [cam], [m], o, with_state=False)
    got, st = _trace(acc, "sirius_hip_trace_batch_noalloc", [cam], [m], o, with_state=False)
    _same(got, ref, o.n_bins)
    assert st["rays"] == st_ref["rays"] == 1920 * 1080
    assert st["steps_accepted"] == st_ref["steps_accepted"] and st["steps_rejected"] == st_ref["steps_rejected"]
    assert (got["cls"] == abi.RAY_DISK).sum() > 0 and (got["cls"] == abi.RAY_ESCAPED).sum() > 0
    # 2: a smaller batch on the scratch the first left behind: 256x256 Schwarzschild, FP64, with
    # the final state requested too
    cam2 = abi.camera(30.0, 80.0, 30.0, 256, 256, disk=(6.0, 20.0))
    m2 = abi.metric_ks(1.0, 0.0)
    o2 = default_opts(abi.PRECISION_FP64)
    o2.n_bins = 8
    print("[case] noalloc 256x256 schwarzschild fp64", flush=True)
    ref2, _ = _trace(acc, "sirius_hip_trace_batch", [cam2], [m2], o2, with_state=True)
    got2, st2 = _trace(acc, "sirius_hip_trace_batch_noalloc", [cam2], [m2], o2, with_state=True)
    _same(got2, ref2, o2.n_bins)
    assert st2["rays"] == 256 * 256
    # and the first batch again after the second: still the same results
    print("[case] noalloc 1920x1080 again", flush=True)
    got3, _ = _trace(acc, "sirius_hip_trace_batch_noalloc", [cam], [m], o, with_state=False)
    _same(got3, ref, o.n_bins)


def test_noalloc_rejects_what_does_not_fit(own_acc):
    acc = own_acc
    big = abi.camera(50.0, 75.0, 20.0, 8192, 8192, disk=(0.0, 30.0))
    m = abi.metric_ks(1.0, 0.9)
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    ro = abi.RayOutput(on_device=1)  # nothing supplied: every array would live in the scratch
    st = abi.Stats()
    ca = (abi.Camera * 1)(big)
    ma = (abi.Metric * 1)(m)
    rc = acc._L.sirius_hip_trace_batch_noalloc(acc._h, ca, 1, ma, 1, C.byref(o), C.byref(ro), C.byref(st))
    assert rc == abi.E_INVALID and "no-allocation scratch" in acc.getLastError()
    host = abi.RayOutput(on_device=0)
    rc = acc._L.sirius_hip_trace_batch_noalloc(acc._h, ca, 1, ma, 1, C.byref(o), C.byref(host), C.byref(st))
    assert rc == abi.E_INVALID
    o.use_surrogate = 1
    small = (abi.Camera * 1)(abi.camera(50.0, 75.0, 20.0, 64, 36))
    rc = acc._L.sirius_hip_trace_batch_noalloc(acc._h, small, 1, ma, 1, C.byref(o), C.byref(ro), C.byref(st))
    assert rc == abi.E_UNSUPPORTED
    # the handle still traces normally afterwards (ordinary allocating mode)
    o.use_surrogate = 0
    got, _ = acc.trace_batch([abi.camera(50.0, 75.0, 20.0, 64, 36)], [m], o)
    assert (got.cls == abi.RAY_DISK).sum() > 0
