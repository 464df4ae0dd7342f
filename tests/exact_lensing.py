"""Escape directions of Schwarzschild rays without an integrator: the orbit equation
(du/dphi)^2 = 1/b^2 - u^2 + 2 u^3 (M = 1) integrated by quadrature from the camera's initial
state to r = 1000, where the direction is that of k^i."""
import numpy as np
from scipy.integrate import quad

from oracle import oracle_py as O


def schwarzschild_escape_dir(cam, m, px, py, r_end=1000.0):
    """Unit escape direction of pixel (px, py) (at the image offset set with O.set_subpixel), or
    None for a ray that does not escape (b below the critical value)."""
    x, p = O.camera_state(cam, m, px, py)
    pos, mom = x[1:], p[1:]
    E = -p[0]
    Lvec = np.cross(pos, mom)
    L = np.linalg.norm(Lvec)
    b = L / E
    r0 = np.linalg.norm(pos)
    e1 = pos / r0
    e2 = np.cross(Lvec / L, e1)
    G = lambda u: 1 / b ** 2 - u * u + 2 * u ** 3
    u0, ue = 1 / r0, 1 / r_end
    k = O.metric(m, x)[1] @ p
    if k[1:] @ e1 < 0:
        # G(u) = 2 (umax - u)(u - ua)(uc - u) with roots ua < 0 < umax < uc (b > b_c);
        # u = umax - s^2 turns the turning-point singularity into a smooth integrand
        roots = np.sort(np.roots([2.0, -1.0, 0.0, 1.0 / b ** 2]).real)
        ua, umax, uc = roots
        if not (ua < 0 < u0 < umax < uc):
            return None
        f = lambda s: 2.0 / np.sqrt(2.0 * (umax - s * s - ua) * (uc - umax + s * s))
        dphi = quad(f, 0, np.sqrt(umax - u0), epsabs=1e-14, epsrel=1e-13, limit=200)[0] + \
            quad(f, 0, np.sqrt(umax - ue), epsabs=1e-14, epsrel=1e-13, limit=200)[0]
    else:
        dphi = quad(lambda u: 1 / np.sqrt(G(u)), ue, u0, epsabs=1e-14, epsrel=1e-13)[0]
    er = np.cos(dphi) * e1 + np.sin(dphi) * e2
    ephi = -np.sin(dphi) * e1 + np.cos(dphi) * e2
    vr = np.sqrt(max(1 - (1 - 2 / r_end) * b * b / r_end ** 2, 0.0))
    d = vr * er + (b / r_end) * ephi
    return d / np.linalg.norm(d)


def angles(d):
    return np.arccos(np.clip(d[2], -1.0, 1.0)), np.arctan2(d[1], d[0])


def schwarzschild_jacobian(cam, m, px, py, eps=0.01):
    """d(theta, phi)/d(u, v) of the exact escape direction by central differences over +-eps pixel:
    (da/du, da/dv, db/du, db/dv), or None if an offset ray does not escape."""
    J = np.zeros(4)
    for a in (0, 1):
        vals = []
        for sg in (1.0, -1.0):
            O.set_subpixel(sg * eps if a == 0 else 0.0, sg * eps if a == 1 else 0.0)
            try:
                d = schwarzschild_escape_dir(cam, m, px, py)
            finally:
                O.set_subpixel(0.0, 0.0)
            if d is None:
                return None
            vals.append(angles(d))
        J[a] = (vals[0][0] - vals[1][0]) / (2 * eps)
        J[2 + a] = (np.remainder(vals[0][1] - vals[1][1] + np.pi, 2 * np.pi) - np.pi) / (2 * eps)
    return J
