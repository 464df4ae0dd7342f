"""Restatement of the launch framebuffer's shading (numpy): the disk colour chain of
RenderSession::shadeDiskHit (SNRS001A.cpp:315-434) for a thin disk's first crossing, with its
helpers Spectral::blackbodyToRGB (PHSP001A.h:181-205), ColorModes::applyColorMode TrueColor
(PHSC001A.h:343-366) and Spectral::applyLimbDarkening (PHSP001A.h:306-322), and the pixel grid
average of shadePixel (SNRS001A.cpp:480-560).

TEST INFRASTRUCTURE ONLY: imported by tests/ (the checker of the shading kernel in
sirius_amd/csrc/util_kernels.hip). Pinned against the reference's own functions, compiled in
place by oracle/Makefile (tests/golden/ref_shading.npz, tests/test_oracle_shading.py).

Arithmetic widths follow the reference: the spectrum in FP64, colour sums and products in
float32, one rounding per operation.
"""
import numpy as np

_H, _C, _K = 6.62607015e-34, 2.99792458e8, 1.380649e-23
_C1, _C2 = 2.0 * _H * _C * _C, _H * _C / _K
_DL = (780e-9 - 380e-9) / 32
_LAMBDA = 380e-9 + (np.arange(32) + 0.5) * _DL
f32 = np.float32


def _lobe(nm, a, mu, s_lo, s_hi):
    t = (nm - mu) * np.where(nm < mu, s_lo, s_hi)
    return a * np.exp(-0.5 * t * t)


def _cmf(nm):
    """CIE 1931 matching functions, the piecewise-Gaussian fit of Wyman, Sloan and Shirley
    (JCGT 2013), each rounded to float32."""
    x = _lobe(nm, 0.362, 442.0, 0.0624, 0.0374) + _lobe(nm, 1.056, 599.8, 0.0264, 0.0323) - \
        _lobe(nm, 0.065, 501.1, 0.0490, 0.0382)
    y = _lobe(nm, 0.821, 568.8, 0.0213, 0.0247) + _lobe(nm, 0.286, 530.9, 0.0613, 0.0322)
    z = _lobe(nm, 1.217, 437.0, 0.0845, 0.0278) + _lobe(nm, 0.681, 459.0, 0.0385, 0.0725)
    return x.astype(f32), y.astype(f32), z.astype(f32)


def blackbody_rgb(T):
    """Blackbody colour of temperatures T (K) [n] -> float32 [n, 3], brightest channel 1."""
    T = np.atleast_1d(np.asarray(T, np.float64))
    X = np.zeros(len(T), f32)
    Y = np.zeros(len(T), f32)
    Z = np.zeros(len(T), f32)
    for lam in _LAMBDA:
        x = _C2 / (lam * T)
        with np.errstate(over="ignore"):
            B = np.where((T <= 0) | (x > 700), 0.0, _C1 / (lam ** 5 * (np.exp(np.minimum(x, 700.0)) - 1.0)))
        w = (B * _DL).astype(f32)
        sx, sy, sz = _cmf(lam * 1e9)
        X = X + sx * w
        Y = Y + sy * w
        Z = Z + sz * w
    r = f32(3.2404542) * X - f32(1.5371385) * Y - f32(0.4985314) * Z
    g = f32(-0.9692660) * X + f32(1.8760108) * Y + f32(0.0415560) * Z
    b = f32(0.0556434) * X - f32(0.2040259) * Y + f32(1.0572252) * Z
    m = np.maximum(np.maximum(np.maximum(r, g), b), f32(0.001))
    return np.stack([r / m, g / m, b / m], 1).astype(f32)


def true_color(T_emit, g, intensity):
    """applyColorMode(TrueColor): blackbody colour of clamp(T_emit g 30000 K, 1000, 1e5) times
    intensity g^4 -> float32 [n, 3]."""
    T_emit, g, intensity = (np.atleast_1d(np.asarray(v, f32)) for v in (T_emit, g, intensity))
    tk = np.clip((T_emit * g) * f32(30000.0), f32(1000.0), f32(100000.0))
    rgb = blackbody_rgb(tk.astype(np.float64))
    g4 = ((g * g) * g) * g
    return rgb * (intensity * g4)[:, None]


def limb_coefficients():
    """u(lambda) of the linear law, falling linearly from 0.9 at 400 nm to 0.4 at 700 nm, at
    650, 550 and 450 nm (red, green, blue)."""
    nm = np.array([650.0, 550.0, 450.0])
    t = np.clip((nm - 400.0) / 300.0, 0.0, 1.0)
    return 0.9 + t * (0.4 - 0.9)


def limb_darkening(rgb, cos_i):
    """(1 + u cos i) / (1 + u) per channel (zero colour at cos i <= 0) -> float32 [n, 3]."""
    cos_i = np.atleast_1d(np.asarray(cos_i, np.float64))
    u = limb_coefficients()
    fac = ((1.0 + u[None, :] * cos_i[:, None]) / (1.0 + u[None, :])).astype(f32)
    out = np.asarray(rgb, f32) * fac
    return np.where((cos_i > 0)[:, None], out, f32(0.0)).astype(f32)


def disk_radiance(T_e, g, cos_i, boost=5.0):
    """Radiance of thin-disk hits: emitted temperature T_e (K), redshift g, cosine of the
    emission angle from the disk normal -> float32 [n, 3]. The chain of shadeDiskHit for the
    first crossing (order demagnification 1, no motion blur, magnification 1): TrueColor of
    (T_e / 30000 K, g, t^4) with t = (T_e / 30000 K) g, limb darkening, g^4 beaming, boost."""
    te = (np.asarray(T_e, np.float64) / 30000.0).astype(f32)
    g = np.asarray(g, f32)
    t = te * g
    inten = ((t * t) * t) * t
    rgb = true_color(te, g, inten)
    rgb = limb_darkening(rgb, cos_i)
    g4 = ((g * g) * g) * g
    out = (rgb * g4[:, None]) * f32(boost)
    live = (np.asarray(T_e) > 0) & (g > 0)
    return np.where(live[:, None], out, f32(0.0)).astype(f32)


def pixel_average(sub, w, h, gs):
    """Average of each pixel's gs x gs rays: sub [gs h, gs w, 3] -> [h, w, 3] (float32 sums in
    row order, then times 1 / gs^2)."""
    sub = np.asarray(sub, f32).reshape(h, gs, w, gs, 3)
    acc = np.zeros((h, w, 3), f32)
    for sy in range(gs):
        for sx in range(gs):
            acc = acc + sub[:, sy, :, sx, :]
    return acc * f32(1.0 / (gs * gs))
