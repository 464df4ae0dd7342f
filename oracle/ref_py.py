"""ctypes binding of oracle/_ref/libsirius_ref.so (the reference's own CPU code, compiled in
place by `make -C oracle ref`). TEST INFRASTRUCTURE ONLY: used to make golden vectors and as
the CPU baseline; the library exists only where the reference tree was present at build time.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libsirius_ref.so")
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB_PATH)
        d, dp, i, ip = C.c_double, C.POINTER(C.c_double), C.c_int, C.c_void_p
        L.ref_ks_evaluate.argtypes = [d, d, d, d, dp, dp, dp]
        L.ref_mt_evaluate.argtypes = [d, dp, dp, dp]
        L.ref_warp_evaluate.argtypes = [d, d, d, dp, dp, dp]
        L.ref_ks_radius.restype = d
        L.ref_ks_radius.argtypes = [d, d, d, d]
        L.ref_camera_ray.argtypes = [d, d, d, d, i, i, i, i, dp, dp]
        L.ref_init_state.argtypes = [d, d, d, d, d, d, d, d, i, i, i, i, dp, dp]
        L.ref_normalize_null.restype = d
        L.ref_normalize_null.argtypes = [dp, dp]
        L.ref_rk45_step.restype = i
        L.ref_rk45_step.argtypes = [d, d, d, d, dp, dp, dp, d, d, d, d]
        L.ref_trace.argtypes = [d, d, d, d, d, d, d, d, i, i, ip, C.c_int64, d, d, d, d, d, i,
                                d, d, d, d, i, i, ip, ip, ip, ip, ip, ip]
        L.ref_isco.restype = d
        L.ref_isco.argtypes = [d]
        L.ref_pt_q.restype = d
        L.ref_pt_q.argtypes = [d, d]
        L.ref_circular.argtypes = [d, d, dp, dp, dp]
        L.ref_symplectic_run.argtypes = [d, d, d, d, d, d, d, i, d, dp, dp, dp]
        f, u32 = C.c_float, C.c_uint32
        L.ref_starfield_generate.restype = i
        L.ref_starfield_generate.argtypes = [u32, f, f, f, u32, ip, i]
        L.ref_starfield_sample.argtypes = [f, f, f, i, ip, i, ip, i, ip]
        L.ref_blackbody_rgb.argtypes = [ip, i, ip]
        L.ref_true_color.argtypes = [ip, ip, ip, i, ip]
        L.ref_limb_darkening.argtypes = [ip, ip, i, ip]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ks_evaluate(M, a, Q, Lam, x):
    g, dg = np.zeros(16), np.zeros(64)
    x = np.ascontiguousarray(x, np.float64)
    lib().ref_ks_evaluate(M, a, Q, Lam, _p(x), _p(g), _p(dg))
    return g.reshape(4, 4), dg.reshape(4, 4, 4)


def mt_evaluate(b0, x):
    g, dg = np.zeros(16), np.zeros(64)
    x = np.ascontiguousarray(x, np.float64)
    lib().ref_mt_evaluate(b0, _p(x), _p(g), _p(dg))
    return g.reshape(4, 4), dg.reshape(4, 4, 4)


def warp_evaluate(v, R, sigma, x):
    g, dg = np.zeros(16), np.zeros(64)
    x = np.ascontiguousarray(x, np.float64)
    lib().ref_warp_evaluate(v, R, sigma, _p(x), _p(g), _p(dg))
    return g.reshape(4, 4), dg.reshape(4, 4, 4)


def camera_ray(r, theta, phi, fov, w, h, px, py):
    o, d = np.zeros(4), np.zeros(4)
    lib().ref_camera_ray(r, theta, phi, fov, w, h, px, py, _p(o), _p(d))
    return o, d


def init_state(M, a, Q, Lam, r, theta, phi, fov, w, h, px, py):
    x, k = np.zeros(4), np.zeros(4)
    lib().ref_init_state(M, a, Q, Lam, r, theta, phi, fov, w, h, px, py, _p(x), _p(k))
    return x, k


def trace(M, a, Q, Lam, r_obs, theta, phi, fov, w, h, pix, rtol=1e-8, atol=1e-10, h0=0.01,
          hmin=1e-8, hmax=2.0, max_steps=20000, r_cap=2.02, r_esc=1000.0, disk=(6.0, 20.0),
          use_disk=True, threads=0):
    pix = np.ascontiguousarray(pix, np.int64)
    n = len(pix)
    cls = np.zeros(n, np.int32)
    steps = np.zeros(n, np.int32)
    xf, kf = np.zeros((n, 4)), np.zeros((n, 4))
    dr, lam = np.zeros(n), np.zeros(n)
    lib().ref_trace(M, a, Q, Lam, r_obs, theta, phi, fov, w, h, pix.ctypes.data, n, rtol, atol,
                    h0, hmin, hmax, max_steps, r_cap, r_esc, disk[0], disk[1], int(use_disk),
                    threads, cls.ctypes.data, steps.ctypes.data, xf.ctypes.data, kf.ctypes.data,
                    dr.ctypes.data, lam.ctypes.data)
    return dict(cls=cls, steps=steps, x=xf, k=kf, disk_r=dr, lam=lam)


def isco(a):
    return lib().ref_isco(a)


def pt_q(a, r):
    return lib().ref_pt_q(a, r)


def circular(a, r):
    E, L, Om = C.c_double(), C.c_double(), C.c_double()
    lib().ref_circular(a, r, C.byref(E), C.byref(L), C.byref(Om))
    return E.value, L.value, Om.value


def symplectic_run(a, r, th, pr, pth, E, Lz, n, h):
    mh, de, dl = C.c_double(), C.c_double(), C.c_double()
    lib().ref_symplectic_run(a, r, th, pr, pth, E, Lz, n, h, C.byref(mh), C.byref(de), C.byref(dl))
    return mh.value, de.value, dl.value


def starfield_generate(count, dmin, dmax, maglim, seed):
    """StarfieldGenerator::generate: (n, 8) float32 rows in StarEntry order."""
    cap = int(count)
    out = np.zeros((max(cap, 1), 8), np.float32)
    n = lib().ref_starfield_generate(count, dmin, dmax, maglim, seed, out.ctypes.data, cap)
    return out[:n].copy()


def starfield_sample(aperture_mm, focus_pc, brightness, dof, stars, dirs):
    """StarfieldGenerator::sampleStarfield at each unit direction (observer at the origin)."""
    stars = np.ascontiguousarray(stars, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    rgb = np.zeros((len(dirs), 3), np.float32)
    lib().ref_starfield_sample(aperture_mm, focus_pc, brightness, int(dof), stars.ctypes.data, len(stars),
                               dirs.ctypes.data, len(dirs), rgb.ctypes.data)
    return rgb


_img = None
IMAGE_LIB_PATH = os.path.join(_HERE, "_ref", "libsirius_ref_image.so")


def image_lib():
    """The reference's image writers (PNGWriter, EXRWriter, BufferWriter) and the decoders
    vendored with it (stb_image, tinyexr's loader)."""
    global _img
    if _img is None:
        L = C.CDLL(IMAGE_LIB_PATH)
        s, i, p, f = C.c_char_p, C.c_int, C.c_void_p, C.c_float
        L.ref_write_png_rgba.argtypes = [s, i, i, p]
        L.ref_write_exr_rgba.argtypes = [s, i, i, p]
        L.ref_write_ppm_aces.argtypes = [s, p, i, i, f]
        L.ref_write_ppm_srgb.argtypes = [s, i, i, p]
        L.ref_write_pfm.argtypes = [s, i, i, p]
        L.ref_write_raw.argtypes = [s, p, i, i, i]
        L.ref_load_png.argtypes = [C.c_char_p, i, C.POINTER(i), C.POINTER(i), p, i]
        L.ref_load_exr.argtypes = [C.c_char_p, i, C.POINTER(i), C.POINTER(i), p, i]
        _img = L
    return _img


def _written(fn, *args):
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out")
        assert fn(path.encode(), *args) == 1, fn.__name__
        with open(path, "rb") as fh:
            return fh.read()


def write_png(rgba):
    a = np.ascontiguousarray(rgba, np.float32)
    return _written(image_lib().ref_write_png_rgba, a.shape[1], a.shape[0], a.ctypes.data)


def write_exr(rgba):
    a = np.ascontiguousarray(rgba, np.float32)
    return _written(image_lib().ref_write_exr_rgba, a.shape[1], a.shape[0], a.ctypes.data)


def write_ppm_aces(rgba, exposure=1.0):
    a = np.ascontiguousarray(rgba, np.float32)
    L = image_lib()
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out")
        assert L.ref_write_ppm_aces(path.encode(), a.ctypes.data, a.shape[1], a.shape[0], exposure) == 1
        with open(path, "rb") as fh:
            return fh.read()


def write_ppm_srgb(rgb):
    a = np.ascontiguousarray(rgb, np.float32)
    return _written(image_lib().ref_write_ppm_srgb, a.shape[1], a.shape[0], a.ctypes.data)


def write_pfm(rgb):
    a = np.ascontiguousarray(rgb, np.float32)
    return _written(image_lib().ref_write_pfm, a.shape[1], a.shape[0], a.ctypes.data)


def write_raw(rgba):
    a = np.ascontiguousarray(rgba, np.float32)
    L = image_lib()
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out")
        assert L.ref_write_raw(path.encode(), a.ctypes.data, a.shape[1], a.shape[0], a.shape[2]) == 1
        with open(path, "rb") as fh:
            return fh.read()


def load_png(data, h, w):
    out = np.zeros((h, w, 4), np.uint8)
    wi, hi = C.c_int(), C.c_int()
    assert image_lib().ref_load_png(data, len(data), C.byref(wi), C.byref(hi), out.ctypes.data, out.nbytes) == 1
    assert (wi.value, hi.value) == (w, h)
    return out


def load_exr(data, h, w):
    out = np.zeros((h, w, 4), np.float32)
    wi, hi = C.c_int(), C.c_int()
    assert image_lib().ref_load_exr(data, len(data), C.byref(wi), C.byref(hi), out.ctypes.data, out.size) == 1
    assert (wi.value, hi.value) == (w, h)
    return out


def blackbody_rgb(T):
    """Spectral::blackbodyToRGB (PHSP001A.h:181-205) at temperatures T (K): float32 [n, 3]."""
    T = np.ascontiguousarray(T, np.float64)
    out = np.zeros((len(T), 3), np.float32)
    lib().ref_blackbody_rgb(T.ctypes.data, len(T), out.ctypes.data)
    return out


def true_color(T_emit, g, intensity):
    """ColorModes::applyColorMode(TrueColor, ...) (PHSC001A.h:343-366): float32 [n, 3]."""
    T_emit, g, intensity = (np.ascontiguousarray(v, np.float32) for v in (T_emit, g, intensity))
    out = np.zeros((len(T_emit), 3), np.float32)
    lib().ref_true_color(T_emit.ctypes.data, g.ctypes.data, intensity.ctypes.data, len(T_emit), out.ctypes.data)
    return out


def limb_darkening(rgb, cos_i):
    """Spectral::applyLimbDarkening (PHSP001A.h:306-322): float32 [n, 3]."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    cos_i = np.ascontiguousarray(cos_i, np.float64)
    out = np.zeros((len(cos_i), 3), np.float32)
    lib().ref_limb_darkening(rgb.ctypes.data, cos_i.ctypes.data, len(cos_i), out.ctypes.data)
    return out
