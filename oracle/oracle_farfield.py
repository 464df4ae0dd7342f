"""FP64 restatement of the surrogate's analytic far-field reference (numpy, vectorised).

TEST INFRASTRUCTURE ONLY: imported by tests/ (the checker of sirius_amd/csrc/farfield.hpp).

The Schwarzschild (M = 1) null geodesic through a ray's initial position x0 and unit coordinate
direction n stays in their plane; with u = 1/r and b = L/E,
    (du/dphi)^2 = P(u) = 1/b^2 - u^2 + 2 u^3 = 2 (u - u1)(u - u2)(u - u3),  u1 < 0 < u2 < u3,
and an inbound ray sweeps I(ua) = int_{ua}^{u2} du / sqrt(P) from radius 1/ua to its turning
point. The orbit equation is the textbook one (e.g. Chandrasekhar, The Mathematical Theory of
Black Holes, sec. 20); Kerr-Schild coordinates with a = 0 are Schwarzschild's in space, so at
a = 0 this is the integrator's own escape direction (the pin in tests/test_oracle_farfield.py).
The quadrature here is an independent one: the substitution u = u2 - sigma^2 and a 48-point
Gauss-Legendre rule (the device uses 16 points).
"""
import math

import numpy as np

_T, _W = np.polynomial.legendre.leggauss(48)
_T = 0.5 * (_T + 1.0)
_W = 0.5 * _W


def _roots(b):
    c = 1.0 / (b * b)
    th = np.arccos(np.clip(1.0 - 54.0 * c, -1.0, 1.0)) / 3.0
    return (np.cos(th - 4.0 * math.pi / 3.0) / 3.0 + 1.0 / 6.0, np.cos(th - 2.0 * math.pi / 3.0) / 3.0 + 1.0 / 6.0,
            np.cos(th) / 3.0 + 1.0 / 6.0)


def _g(u, u1, u3):
    return 1.0 / np.sqrt(np.maximum(2.0 * (u - u1) * (u3 - u), 1e-300))


def _sweep_sigma(sig, u1, u2, u3):
    u = u2[:, None] - (sig * sig)[:, None] * _T[None, :] ** 2
    return 2.0 * sig * (_g(u, u1[:, None], u3[:, None]) * _W[None, :]).sum(1)


SPIN_SHIFT = 1.1


def spin_shifted_start(x0, a):
    """The reference orbit's start for spin a: x0 displaced by SPIN_SHIFT a along the observer's
    azimuthal unit vector e_phi = z x x0 / |z x x0| (frame dragging's first-order effect on a
    distant ray is a displacement of the lens by a along -e_phi)."""
    x0 = np.asarray(x0, np.float64)
    ephi = np.stack([-x0[:, 1], x0[:, 0], np.zeros(len(x0))], 1)
    ephi /= np.linalg.norm(ephi, axis=1, keepdims=True)
    return x0 + SPIN_SHIFT * np.asarray(a, np.float64).reshape(-1, 1) * ephi


def orbit(x0, n, r_exit=1000.0):
    """Orbit constants of rays (x0, n: arrays [N, 3]): dict of the plane basis rh, et, b, roots,
    sweeps phi_in (start to turning point) and phi_out (turning point to r_exit), and ok (b above
    the critical sqrt(27) and inbound at the start)."""
    x0 = np.asarray(x0, np.float64)
    n = np.asarray(n, np.float64)
    r0 = np.linalg.norm(x0, axis=1)
    rh = x0 / r0[:, None]
    nr = (n * rh).sum(1)
    et = n - nr[:, None] * rh
    nt = np.linalg.norm(et, axis=1)
    et = et / np.maximum(nt, 1e-300)[:, None]
    b = r0 * nt / np.sqrt(nr * nr + (1.0 - 2.0 / r0) * nt * nt)
    ok = (nr < 0) & (b > math.sqrt(27.0)) & (nt > 0)
    u1, u2, u3 = _roots(np.where(ok, b, 100.0))
    phi_in = _sweep_sigma(np.sqrt(np.maximum(u2 - 1.0 / r0, 0.0)), u1, u2, u3)
    phi_out = _sweep_sigma(np.sqrt(np.maximum(u2 - 1.0 / r_exit, 0.0)), u1, u2, u3)
    return dict(rh=rh, et=et, b=b, u1=u1, u2=u2, u3=u3, r0=r0, r_exit=r_exit, phi_in=phi_in, phi_out=phi_out, ok=ok)


def exit_direction(o):
    """Unit coordinate direction at the exit sphere, [N, 3]."""
    Phi = o["phi_in"] + o["phi_out"]
    ue = 1.0 / o["r_exit"]
    re = np.cos(Phi)[:, None] * o["rh"] + np.sin(Phi)[:, None] * o["et"]
    pe = -np.sin(Phi)[:, None] * o["rh"] + np.cos(Phi)[:, None] * o["et"]
    vr = np.sqrt(np.maximum(1.0 - o["b"] ** 2 * ue * ue * (1.0 - 2.0 * ue), 0.0))
    d = vr[:, None] * re + (o["b"] * ue)[:, None] * pe
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def u_at(o, phi):
    """1 / r of the orbit at in-plane azimuth phi from the start (Newton's method on the sweep)."""
    u1, u2, u3 = o["u1"], o["u2"], o["u3"]
    inb = phi <= o["phi_in"]
    target = np.where(inb, o["phi_in"] - phi, phi - o["phi_in"])
    ua = np.where(inb, 1.0 / o["r0"], 1.0 / o["r_exit"])
    smax = np.sqrt(np.maximum(u2 - ua, 0.0))
    sig = target / (2.0 * _g(u2, u1, u3))
    for _ in range(30):
        sig = np.minimum(np.maximum(sig, 0.0), smax)
        sig = sig - (_sweep_sigma(sig, u1, u2, u3) - target) / (2.0 * _g(u2 - sig * sig, u1, u3))
    sig = np.minimum(np.maximum(sig, 0.0), smax)
    return u2 - sig * sig


def crossing_radii(o):
    """Radii of the orbit's crossings of the equatorial plane z = 0 (NaN where there are fewer),
    [N, 3]: the plane meets the orbit's plane along the line of nodes, at in-plane azimuths
    phi_n + k pi inside the sweep."""
    rz, ez = o["rh"][:, 2], o["et"][:, 2]
    phin = np.arctan2(-rz, ez)
    phin = np.where(phin < 0, phin + math.pi, phin)
    Phi = o["phi_in"] + o["phi_out"]
    out = np.full((len(rz), 3), np.nan)
    for k in range(3):
        phi = phin + k * math.pi
        valid = o["ok"] & (phi < Phi)
        r = 1.0 / u_at(o, np.where(valid, phi, 0.0))
        out[:, k] = np.where(valid, r, np.nan)
    return out
