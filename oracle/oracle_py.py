"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry and bench.py's CPU-baseline
leg, never by the sirius_amd package.
"""
import ctypes as C
import os

import numpy as np

from sirius_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libsirius_ref.so")
_lib = None


def build(force=False):
    """Compile the oracle with gcc (OpenMP). Returns the library path."""
    import subprocess
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-C", _HERE, "-s", "liboracle.so"])
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.oracle_metric.argtypes = [C.POINTER(abi.Metric), dp, dp, dp, dp]
        L.oracle_kerr_radius.restype = C.c_double
        L.oracle_kerr_radius.argtypes = [C.POINTER(abi.Metric), C.c_double, C.c_double, C.c_double]
        L.oracle_horizon.restype = C.c_double
        L.oracle_horizon.argtypes = [C.POINTER(abi.Metric)]
        L.oracle_isco.restype = C.c_double
        L.oracle_isco.argtypes = [C.c_double]
        L.oracle_photon_orbit.restype = C.c_double
        L.oracle_photon_orbit.argtypes = [C.c_double, C.c_int]
        L.oracle_page_thorne_f.restype = C.c_double
        L.oracle_page_thorne_f.argtypes = [C.c_double, C.c_double]
        L.oracle_disk_temperature_ratio.restype = C.c_double
        L.oracle_disk_temperature_ratio.argtypes = [C.c_double] * 4
        L.oracle_disk_tmax_f_over_r.restype = C.c_double
        L.oracle_disk_tmax_f_over_r.argtypes = [C.c_double] * 3
        L.oracle_rhs.restype = C.c_double
        L.oracle_rhs.argtypes = [C.POINTER(abi.Metric), dp, dp, dp, dp]
        L.oracle_camera_state.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Metric), C.c_int,
                                          C.c_int, dp, dp]
        L.oracle_critical_distance.restype = C.c_double
        L.oracle_critical_distance.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Metric),
                                               C.c_int, C.c_int]
        L.oracle_bin_centres.argtypes = [C.c_int, C.c_double, C.c_double, dp]
        L.oracle_default_opts.argtypes = [C.POINTER(abi.TraceOpts), C.c_int]
        L.oracle_trace.restype = C.c_int
        L.oracle_trace.argtypes = [C.POINTER(abi.Camera), C.c_int, C.POINTER(abi.Metric), C.c_int,
                                   C.POINTER(abi.TraceOpts), C.c_void_p, C.c_int64,
                                   C.POINTER(abi.RayOutput), C.c_int, C.POINTER(abi.Stats)]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_opts(precision=abi.PRECISION_FP64):
    o = abi.TraceOpts()
    lib().oracle_default_opts(C.byref(o), precision)
    return o


def metric(m, x):
    g, gi, dg = np.zeros(16), np.zeros(16), np.zeros(64)
    x = np.ascontiguousarray(x, np.float64)
    lib().oracle_metric(C.byref(m), _d(x), _d(g), _d(gi), _d(dg))
    return g.reshape(4, 4), gi.reshape(4, 4), dg.reshape(4, 4, 4)


def rhs(m, x, p):
    x = np.ascontiguousarray(x, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    dx, dp = np.zeros(4), np.zeros(4)
    pk = lib().oracle_rhs(C.byref(m), _d(x), _d(p), _d(dx), _d(dp))
    return dx, dp, pk


def camera_state(cam, m, px, py):
    x, p = np.zeros(4), np.zeros(4)
    lib().oracle_camera_state(C.byref(cam), C.byref(m), px, py, _d(x), _d(p))
    return x, p


def critical_distance(cam, m, px, py):
    return lib().oracle_critical_distance(C.byref(cam), C.byref(m), px, py)


def bin_centres(n, x_min=0.1, x_max=10.0):
    c = np.zeros(n)
    lib().oracle_bin_centres(n, x_min, x_max, _d(c))
    return c


def trace(cams, metrics, opts, ray_ids=None, threads=0, with_state=False):
    """Trace the listed rays (all rays when ray_ids is None). Returns (HostRays, stats dict)."""
    cams = list(cams)
    metrics = list(metrics)
    ca = (abi.Camera * len(cams))(*cams)
    ma = (abi.Metric * len(metrics))(*metrics)
    total = sum(c.width * c.height for c in cams)
    if ray_ids is None:
        n, ids_ptr = total, None
    else:
        ids = np.ascontiguousarray(ray_ids, np.int64)
        n, ids_ptr = len(ids), ids.ctypes.data
    out = abi.HostRays(n, opts.n_bins, with_state)
    ro = out.output()
    st = abi.Stats()
    rc = lib().oracle_trace(ca, len(cams), ma, len(metrics), C.byref(opts), ids_ptr, n,
                            C.byref(ro), threads, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle_trace failed: {rc}")
    return out, st.as_dict()


def _mlp_lib():
    L = lib()
    if not getattr(L, "_mlp_ready", False):
        fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.oracle_mlp_param_count.restype = C.c_size_t
        L.oracle_mlp_forward.argtypes = [fp, fp, C.c_int, dp, C.c_int]
        L.oracle_mlp_loss_grad.restype = C.c_double
        L.oracle_mlp_loss_grad.argtypes = [fp, fp, C.POINTER(C.c_int32), fp, C.c_int, dp, C.c_int]
        L._mlp_ready = True
    return L


def mlp_param_count():
    return int(_mlp_lib().oracle_mlp_param_count())


def mlp_forward(params, x, round_bf16=False):
    L = _mlp_lib()
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros((len(x), 6))
    L.oracle_mlp_forward(params.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)),
                         len(x), _d(out), int(round_bf16))
    return out


def mlp_loss_grad(params, x, cls, dirs, round_bf16=False):
    L = _mlp_lib()
    params = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    cls = np.ascontiguousarray(cls, np.int32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    g = np.zeros(len(params))
    loss = L.oracle_mlp_loss_grad(params.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)),
                                  cls.ctypes.data_as(C.POINTER(C.c_int32)), dirs.ctypes.data_as(C.POINTER(C.c_float)),
                                  len(x), _d(g), int(round_bf16))
    return loss, g
