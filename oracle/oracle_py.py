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


# ------------------------------------------------------------------------------------------
# Star field (celestial spheres): restatement of StarfieldGenerator::generate / generateStar
# (src/Sirius.Core/Environment/PHSF001A.h:133-151, 206-249) and ::sampleStarfield
# (PHSF001A.h:153-201). The generator draws from a 32-bit Mersenne twister (std::mt19937, from
# its published definition) through libstdc++'s float canonical draw (one 32-bit word, rounded
# to float, divided by 2^32, kept below 1), and evaluates its formulas in float with the C
# library's float functions; the sampler evaluates the point-spread sum in float32 (numpy),
# products and sums unfused, in the reference's order.
# ------------------------------------------------------------------------------------------
_f32 = np.float32
_libm = None


def _m():
    global _libm
    if _libm is None:
        _libm = C.CDLL("libm.so.6")
        for name in ("acosf", "sinf", "cosf", "cbrtf", "log10f", "sqrtf"):
            getattr(_libm, name).restype = C.c_float
            getattr(_libm, name).argtypes = [C.c_float]
        _libm.powf.restype = C.c_float
        _libm.powf.argtypes = [C.c_float, C.c_float]
    return _libm


class MT19937:
    """32-bit Mersenne twister (Matsumoto & Nishimura 1998), the std::mt19937 engine."""

    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.i = 624

    def next(self):
        if self.i >= 624:
            for k in range(624):
                y = (self.mt[k] & 0x80000000) | (self.mt[(k + 1) % 624] & 0x7FFFFFFF)
                self.mt[k] = self.mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF

    def uniform_float(self):
        """std::uniform_real_distribution<float>(0, 1) over this engine (libstdc++)."""
        r = _f32(self.next()) / _f32(4294967296.0)
        return r if r < _f32(1.0) else np.nextafter(_f32(1.0), _f32(0.0))


def _validate_starfield(count, dmin, dmax, maglim):
    dmin = max(_f32(dmin), _f32(0.1))
    dmax = max(_f32(dmax), _f32(dmin + _f32(1.0)))
    return min(int(count), 10000000), dmin, dmax, min(max(_f32(maglim), _f32(0.0)), _f32(20.0))


def starfield_generate(count, dmin, dmax, maglim, seed):
    """The catalogue: (n, 8) float32 rows (direction, distance, magnitude, B-V, temperature, 0)."""
    m = _m()
    count, dmin, dmax, maglim = _validate_starfield(count, dmin, dmax, maglim)
    g = MT19937(seed)
    U = g.uniform_float
    pi = _f32(np.pi)
    rows = []
    for _ in range(count):
        th = _f32(m.acosf(_f32(2.0) * U() - _f32(1.0)))
        ph = _f32(2.0) * pi * U()
        st, ct = _f32(m.sinf(th)), _f32(m.cosf(th))
        d = (st * _f32(m.cosf(ph)), st * _f32(m.sinf(ph)), ct)
        u = U()
        dist = _f32(m.cbrtf(u * (dmax * dmax * dmax) + (_f32(1.0) - u) * (dmin * dmin * dmin)))
        absm = _f32(-2.0) + _f32(12.0) * _f32(m.powf(U(), _f32(0.3)))
        mag = absm + _f32(5.0) * _f32(m.log10f(dist / _f32(10.0)))
        t = (absm + _f32(2.0)) / _f32(12.0)
        bv = min(max(_f32(-0.3) + _f32(2.0) * t + _f32(0.3) * (U() - _f32(0.5)), _f32(-0.4)), _f32(2.0))
        temp = _f32(4600.0) / (_f32(0.92) + bv) if bv > _f32(-0.1) else _f32(30000.0) + _f32(10000.0) * (-bv - _f32(0.1))
        temp = min(max(temp, _f32(2500.0)), _f32(50000.0))
        if mag <= maglim:
            rows.append((d[0], d[1], d[2], dist, mag, bv, temp, _f32(0.0)))
    return np.array(rows, np.float32).reshape(-1, 8)


def _star_colour(T):
    m = _m()
    T = _f32(T) if T > 0 else _f32(5778.0)
    tn = T / _f32(6500.0)
    if T <= _f32(6500.0):
        c = [_f32(1.0), _f32(m.powf(tn, _f32(0.4))), _f32(m.powf(tn, _f32(0.8)))]
    else:
        c = [_f32(m.powf(_f32(1.0) / tn, _f32(0.3))), _f32(m.powf(_f32(1.0) / tn, _f32(0.15))), _f32(1.0)]
    mx = max(c)
    return [x / mx for x in c]


def starfield_sample(stars, dirs, aperture_mm=50.0, focus_pc=100.0, brightness=1.0, dof=True):
    """Star light (n, 3) at unit directions dirs (n, 3), observer at the origin."""
    m = _m()
    stars = np.asarray(stars, np.float32)
    dirs = np.asarray(dirs, np.float32)
    out = np.zeros((len(dirs), 3), np.float32)
    b = _f32(brightness)
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        for s in stars:
            dx, dy, dz = s[0] * s[3] - _f32(0.0), s[1] * s[3] - _f32(0.0), s[2] * s[3] - _f32(0.0)
            d = _f32(m.sqrtf(dx * dx + dy * dy + dz * dz))
            if d < _f32(1e-6):
                continue
            px, py, pz = dx / d, dy / d, dz / d
            sig_dof = _f32(0.0)
            if dof and aperture_mm > 0:
                sig_dof = _f32(aperture_mm) * _f32(abs(_f32(1.0) / s[3] - _f32(1.0) / _f32(focus_pc))) * _f32(1e-3)
            sigma = _f32(m.sqrtf(_f32(1e-6) * _f32(1e-6) + sig_dof * sig_dof))
            c = (dirs[:, 0] * px + dirs[:, 1] * py) + dirs[:, 2] * pz
            ang = np.arccos(np.clip(c, _f32(-1.0), _f32(1.0))).astype(np.float32)
            psf = np.exp((_f32(-0.5) * ang * ang) / (sigma * sigma)).astype(np.float32)
            keep = psf >= _f32(1e-6)
            if not keep.any():
                continue
            inten = (_f32(m.powf(_f32(10.0), _f32(-0.4) * s[4])) * psf) * b
            col = _star_colour(s[6])
            for k in range(3):
                out[keep, k] = out[keep, k] + col[k] * inten[keep]
    return out


# ---- ray bundles: derivatives of traced outcomes with respect to the image point ------------
def set_subpixel(du, dv):
    """Offsets every later ray's image point by (du, dv) pixels (0, 0 restores the pixel centres)."""
    L = lib()
    L.oracle_set_subpixel.argtypes = [C.c_double, C.c_double]
    L.oracle_set_subpixel.restype = None
    L.oracle_set_subpixel(du, dv)


def trace_subpixel(cams, metrics, opts, ray_ids, du, dv):
    set_subpixel(du, dv)
    try:
        return trace(cams, metrics, opts, ray_ids=ray_ids, with_state=True)[0]
    finally:
        set_subpixel(0.0, 0.0)


def outcome_coords(metric_, rays):
    """Per ray, from the final FP64 state: the escape direction (theta, phi) of k^i (escaped rays of
    the Cartesian families) and the position (X, Y) (rays on the disk), as (2, n) each."""
    n = len(rays.cls)
    ang = np.zeros((2, n))
    xy = np.zeros((2, n))
    for i in range(n):
        x, p = rays.state[i, :4], rays.state[i, 4:]
        k = metric(metric_, x)[1] @ p
        d = k[1:] / np.linalg.norm(k[1:])
        ang[0, i] = np.arccos(np.clip(d[2], -1.0, 1.0))
        ang[1, i] = np.arctan2(d[1], d[0])
        xy[:, i] = x[1], x[2]
    return ang, xy


def lens_jacobian(cams, metrics, opts, ray_ids, eps=0.01):
    """Central differences over +-eps pixel of whole FP64 traces: J (4, n) = d(a, b)/d(u, v) with
    (a, b) = (theta, phi) for escaped rays and (X, Y) for disk rays (the bundle outputs' layout),
    the central classes, and a mask of rays whose four offset traces end in the same class."""
    base = trace(cams, metrics, opts, ray_ids=ray_ids, with_state=True)[0]
    m = metrics[cams[0].metric_index]
    J = np.zeros((4, len(ray_ids)))
    same = np.ones(len(ray_ids), bool)
    for a in (0, 1):
        plus = trace_subpixel(cams, metrics, opts, ray_ids, eps if a == 0 else 0.0, eps if a == 1 else 0.0)
        minus = trace_subpixel(cams, metrics, opts, ray_ids, -eps if a == 0 else 0.0, -eps if a == 1 else 0.0)
        same &= (plus.cls == base.cls) & (minus.cls == base.cls)
        (ap, xyp), (am, xym) = outcome_coords(m, plus), outcome_coords(m, minus)
        dph = np.remainder(ap[1] - am[1] + np.pi, 2 * np.pi) - np.pi
        esc = base.cls == 0
        J[a] = np.where(esc, (ap[0] - am[0]) / (2 * eps), (xyp[0] - xym[0]) / (2 * eps))
        J[2 + a] = np.where(esc, dph / (2 * eps), (xyp[1] - xym[1]) / (2 * eps))
    return J, base.cls, same
