"""The restatement of the surrogate's analytic far-field reference (oracle/oracle_farfield.py)
against the integrator. With a = 0 Kerr-Schild space is Schwarzschild space, so the orbit
equation's exit direction must be the oracle's traced escape direction, and the orbit's crossings
of the equatorial plane must sort disk hits from escapes."""
import math

import numpy as np

from oracle import oracle_farfield as F
from oracle import oracle_py as O
from sirius_amd import abi


def _rays(cam, m):
    """Initial position and unit coordinate direction of every pixel's ray (the oracle's camera)."""
    n = cam.width * cam.height
    x0, nv = np.zeros((n, 3)), np.zeros((n, 3))
    for i in range(n):
        x, p = O.camera_state(cam, m, i % cam.width, i // cam.width)
        dx, _, _ = O.rhs(m, x, p)
        x0[i], nv[i] = x[1:], dx[1:] / np.linalg.norm(dx[1:])
    return x0, nv


def _angle(u, v):
    u = u / np.linalg.norm(u, axis=1, keepdims=True)
    v = v / np.linalg.norm(v, axis=1, keepdims=True)
    return 2.0 * np.arcsin(np.minimum(np.linalg.norm(u - v, axis=1) / 2.0, 1.0))


def _crit(cam, m):
    return np.array([O.critical_distance(cam, m, i % cam.width, i // cam.width)
                     for i in range(cam.width * cam.height)])


def test_exit_direction_is_the_integrators_at_zero_spin():
    m = abi.metric_ks(1.0, 0.0)
    cam = abi.camera(50.0, 60.0, 40.0, 32, 32, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP64)
    o.emission, o.n_bins = abi.EMISSION_NONE, 1
    o.rtol, o.atol = 1e-11, 1e-13
    ref, _ = O.trace([cam], [m], o, with_state=True)
    x0, nv = _rays(cam, m)
    orb = F.orbit(x0, nv)
    esc = (ref.cls == abi.RAY_ESCAPED) & (np.abs(_crit(cam, m)) > 0.5)
    assert esc.sum() > 400 and np.all(orb["ok"][esc])
    # the traced direction in FP64: k^i at the final state
    st = ref.state[esc]
    kd = np.array([O.rhs(m, s[:4], s[4:])[0][1:] for s in st])
    err = _angle(F.exit_direction(orb)[esc], kd)
    assert err.max() < 1e-8, err.max()
    # and rays below the critical impact parameter are the captured ones
    cap = ref.cls == abi.RAY_CAPTURED
    assert np.array_equal(~orb["ok"] & (np.abs(_crit(cam, m)) > 0.5), cap & (np.abs(_crit(cam, m)) > 0.5))


def test_equatorial_crossings_sort_disk_hits_from_escapes_at_zero_spin():
    m = abi.metric_ks(1.0, 0.0)
    cam = abi.camera(30.0, 80.0, 30.0, 48, 48, disk=(6.0, 20.0))
    o = O.default_opts(abi.PRECISION_FP64)
    o.n_bins = 1
    ref, _ = O.trace([cam], [m], o)
    x0, nv = _rays(cam, m)
    orb = F.orbit(x0, nv)
    rc = F.crossing_radii(orb)
    hit = np.any((rc >= 6.0) & (rc <= 20.0), axis=1)
    live = orb["ok"] & (np.abs(_crit(cam, m)) > 0.5)
    near_edge = np.any(np.minimum(np.abs(rc - 6.0), np.abs(rc - 20.0)) < 1e-6, axis=1)
    keep = live & ~near_edge
    disk = ref.cls == abi.RAY_DISK
    assert (keep & disk).sum() > 300 and (keep & ~disk).sum() > 100
    assert np.array_equal(hit[keep], disk[keep])
    # a disk ray's first crossing inside the annulus is where the integrator landed
    dk = keep & disk
    first = np.array([next(r for r in row if 6.0 <= r <= 20.0) for row in rc[dk]])
    assert np.allclose(first, ref.disk_r[dk], rtol=1e-6)


def test_orbit_inverse_and_closed_forms():
    # rays from and to infinity: the total deflection phi_in + phi_out - pi against Darwin's exact
    # form (closest approach r0, b^2 = r0^3 / (r0 - 2), Q^2 = (r0 - 2)(r0 + 6)):
    #   delta = -pi + 4 sqrt(r0 / Q) (K(k) - F(zeta, k)), k^2 = (Q - r0 + 6) / (2 Q),
    #   sin^2 zeta = (Q - r0 + 2) / (Q - r0 + 6)
    from scipy.special import ellipk, ellipkinc
    r0 = np.array([4.0, 6.0, 10.0, 30.0, 200.0])
    b = np.sqrt(r0 ** 3 / (r0 - 2.0))
    Q = np.sqrt((r0 - 2.0) * (r0 + 6.0))
    k2 = (Q - r0 + 6.0) / (2.0 * Q)
    zeta = np.arcsin(np.sqrt((Q - r0 + 2.0) / (Q - r0 + 6.0)))
    darwin = -math.pi + 4.0 * np.sqrt(r0 / Q) * (ellipk(k2) - ellipkinc(zeta, k2))
    R = 1e14
    x0 = np.zeros((len(b), 3))
    x0[:, 0] = R
    nv = np.stack([-np.sqrt(1.0 - (b / R) ** 2), b / R, np.zeros(len(b))], 1)
    orb = F.orbit(x0, nv, r_exit=R)
    assert np.allclose(1.0 / orb["u2"], r0, rtol=1e-10)
    assert np.allclose(orb["phi_in"] + orb["phi_out"] - math.pi, darwin, rtol=1e-9, atol=1e-12)
    # u_at inverts the sweep: the azimuth of the turning point is phi_in, of the exit Phi
    cam = abi.camera(50.0, 70.0, 30.0, 8, 8)
    x0, nv = _rays(cam, abi.metric_ks(1.0, 0.0))
    orb = F.orbit(x0, nv)
    ok = orb["ok"]
    assert np.allclose(F.u_at(orb, orb["phi_in"])[ok], orb["u2"][ok], rtol=1e-12)
    assert np.allclose(F.u_at(orb, orb["phi_in"] + orb["phi_out"])[ok], 1e-3, rtol=1e-9)
    assert np.allclose(F.u_at(orb, np.zeros(len(ok)))[ok], 1.0 / orb["r0"][ok], rtol=1e-9)
