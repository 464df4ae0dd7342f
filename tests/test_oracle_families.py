"""Pins of the oracle's traces for the wormhole, the warp bubble and the symplectic integrator
against results fixed independently of it (closed-form quadrature, an independent integrator
of another formulation, convergence order).

Reference files: Morris-Thorne metric src/Sirius.Core/Metric/PHMT101A.h:285-345, Alcubierre
src/Sirius.Core/Metric/PHMT102A.h:485-578, Yoshida composition
src/Sirius.Core/Symplectic/PHSI001A.h:43-80.
"""
import numpy as np
import pytest
from scipy.integrate import quad, solve_ivp
from scipy.special import ellipk, ellipkinc

from oracle import oracle_py as O
from sirius_amd import abi


def _ellis_sweep(L, b0, l0, nr, le):
    """Planar angle swept by an Ellis-wormhole null ray (Phi = 0, r^2 = l^2 + b0^2) from the
    observer at l0 to |l| = le, and the side it leaves on. With E = 1 and p_psi = L:
    dpsi/dl = L / (r^2 sqrt(1 - L^2 / r^2)). Rays with L > b0 heading inward turn at
    r = L; there the substitution r = L / sin(chi) turns each leg into the incomplete
    elliptic integral F(pi/2 | m) - F(asin(L / r) | m), m = (b0 / L)^2."""
    f = lambda l: L / ((l * l + b0 * b0) * np.sqrt(1.0 - L * L / (l * l + b0 * b0)))
    q = dict(epsabs=1e-13, epsrel=1e-13, limit=400)
    if nr >= 0:
        return quad(f, l0, le, **q)[0], 1
    if L < b0:  # through the throat to the other side
        return quad(f, 0, l0, **q)[0] + quad(f, 0, le, **q)[0], -1
    m = (b0 / L) ** 2
    leg = lambda l: ellipk(m) - ellipkinc(np.arcsin(L / np.sqrt(l * l + b0 * b0)), m)
    return leg(l0) + leg(le), 1


def test_wormhole_sweep_angle_matches_elliptic_quadrature():
    b0 = 1.0
    m = abi.metric_mt(b0)
    w, h = 48, 27
    # one camera facing the throat, one facing away (yaw = pi) for outgoing rays
    cams = [abi.camera(10.0, 90.0, 60.0, w, h), abi.camera(10.0, 90.0, 60.0, w, h, yaw=np.pi)]
    o = O.default_opts(abi.PRECISION_FP64)
    out, _ = O.trace(cams, [m], o, with_state=True)
    assert np.all(out.cls == abi.RAY_ESCAPED)
    kinds = {"outward": 0, "through": 0, "turning": 0}
    worst = 0.0
    for i in range(2 * w * h):
        cam, j = cams[i // (w * h)], i % (w * h)
        x, p = O.camera_state(cam, m, j % w, j // w)
        L, l0, nr = p[2], x[1], p[1]
        st = out.state[i]
        assert abs(abs(st[1]) - o.r_escape) < 1e-6
        psi, side = _ellis_sweep(L, b0, l0, nr, abs(st[1]))
        kinds["outward" if nr >= 0 else ("through" if L < b0 else "turning")] += 1
        assert np.sign(st[1]) == side
        assert bool(out.flags[i] & abi.FLAG_FAR_SIDE) == (side < 0)
        worst = max(worst, abs(st[2] - psi))
    assert min(kinds.values()) > 5, kinds
    assert worst < 1e-7, worst


def _bubble(v, R, sigma):
    def f_df(r):
        t1, t2, tr = np.tanh(sigma * (r + R)), np.tanh(sigma * (r - R)), np.tanh(sigma * R)
        return (t1 - t2) / (2 * tr), sigma * ((1 - t1 * t1) - (1 - t2 * t2)) / (2 * tr)

    def ray(t, y):
        """Null rays of ds^2 = -dt^2 + (dx - v f dt)^2 + dy^2 + dz^2 in coordinate time: the
        dispersion relation omega = |k| + v f k_x gives dx/dt = k/|k| + v f e_x and
        dk/dt = -v k_x grad f."""
        x, k = y[:3], y[3:]
        r = np.linalg.norm(x)
        f, df = f_df(r)
        return np.r_[k / np.linalg.norm(k) + np.array([v * f, 0.0, 0.0]), -v * df * k[0] * x / r]
    return ray


@pytest.mark.parametrize("rtol,bar", [(1e-10, 1e-7), (1e-8, 3e-6)])
def test_warp_bubble_rays_match_independent_river_integration(rtol, bar):
    """Escape directions of the oracle (DP45 on Hamilton's equations in the affine parameter)
    against an 8th-order integration of the same spacetime's ray equations written in
    coordinate time from the dispersion relation. Many rays cross the bubble wall and are
    deflected by up to ~2 rad."""
    v, R, sigma = 0.5, 1.0, 8.0
    m = abi.metric_warp(v, R, sigma)
    w, h = 24, 16
    cam = abi.camera(10.0, 90.0, 40.0, w, h, phi_deg=30.0)
    o = O.default_opts(abi.PRECISION_FP64)
    o.rtol, o.atol = rtol, rtol * 1e-2
    out, _ = O.trace([cam], [m], o, with_state=True)
    assert np.all(out.cls == abi.RAY_ESCAPED)
    ray = _bubble(v, R, sigma)
    esc = lambda t, y: np.linalg.norm(y[:3]) - o.r_escape
    esc.terminal, esc.direction = True, 1
    worst, deflected = 0.0, 0
    for i in range(w * h):
        x, p = O.camera_state(cam, m, i % w, i // w)
        sol = solve_ivp(ray, (0, 5000), np.r_[x[1:], p[1:]], method="DOP853", rtol=1e-12, atol=1e-12, events=esc)
        d_ref = ray(0, sol.y[:, -1])[:3]
        d_ref /= np.linalg.norm(d_ref)
        st = out.state[i]
        dx, _, _ = O.rhs(m, st[:4], st[4:])
        d = dx[1:] / np.linalg.norm(dx[1:])
        worst = max(worst, np.arccos(np.clip(d @ d_ref, -1, 1)))
        k0 = p[1:] / np.linalg.norm(p[1:])
        deflected += np.arccos(np.clip(k0 @ d, -1, 1)) > 1e-2
    assert deflected > 20
    assert worst < bar, worst


def test_symplectic_convergence_order_and_agreement_with_dp45():
    """Tao's extended phase space flow with the Yoshida triple jump is 4th order: the null
    constraint error falls ~16x per halving of h (~4x for the plain 2nd-order Strang step).
    A wrong composition weight drops the order. Classes match DP45 and escape directions agree
    to the step error."""
    m = abi.metric_ks(1.0, 0.9)
    w, h = 16, 12
    cam = abi.camera(30.0, 70.0, 30.0, w, h, disk=(0.0, 0.0))
    o = O.default_opts(abi.PRECISION_FP64)
    o.emission = abi.EMISSION_NONE
    o.r_escape = 60.0
    o.max_steps = 200000
    dp, _ = O.trace([cam], [m], o)
    o.integrator = abi.INTEGRATOR_SYMPLECTIC
    for order, lo, hi in ((4, 10.0, 30.0), (2, 3.0, 6.0)):
        o.symplectic_order = order
        drift = []
        for step in (0.04, 0.02, 0.01):
            o.symplectic_h = step
            s, _ = O.trace([cam], [m], o)
            assert np.array_equal(s.cls, dp.cls)
            drift.append(np.median(s.null_drift))
        ratios = np.array(drift[:-1]) / np.array(drift[1:])
        assert np.all((ratios > lo) & (ratios < hi)), (order, ratios)
        if order == 4:
            esc = (s.cls == 0) & (dp.cls == 0)
            u, v = s.direction()[esc], dp.direction()[esc]
            ang = np.arccos(np.clip((u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1), -1, 1))
            assert np.percentile(ang, 90) < 1e-5 and ang.max() < 1e-3, (np.percentile(ang, 90), ang.max())


def test_thin_disk_redshift_and_intensity_from_conserved_quantities():
    """Exact g-factor of a prograde Keplerian emitter in Kerr (Bardeen, Press & Teukolsky 1972):
    u^t = (r^{3/2} + a) / (r^{3/4} sqrt(r^{3/2} - 3 r^{1/2} + 2a)), Omega = 1 / (r^{3/2} + a),
    g = nu_obs / nu_emit = obs_norm / (u^t (1 - Omega L_z / E)), with E = -p_t and
    L_z = x p_y - y p_x conserved along the ray, so they are taken from the camera state,
    independent of the trace. The spectrum is then I_k = x_k^3 / expm1(x_k / (g T(r)/T_max))."""
    a = 0.9
    m = abi.metric_ks(1.0, a)
    w, h = 64, 36
    cam = abi.camera(50.0, 75.0, 20.0, w, h, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP64)
    out, _ = O.trace([cam], [m], o, with_state=True)
    x_obs, _ = O.camera_state(cam, m, 0, 0)
    obs_norm = 1.0 / np.sqrt(-O.metric(m, x_obs)[0][0, 0])
    r_in = O.lib().oracle_isco(a)
    bins = O.bin_centres(o.n_bins, o.x_min, o.x_max)
    disk = np.where(out.cls == abi.RAY_DISK)[0]
    assert len(disk) > 200
    worst_g = worst_i = 0.0
    for i in disk:
        x, p = O.camera_state(cam, m, i % w, i // w)
        E, Lz = -p[0], x[1] * p[2] - x[2] * p[1]
        st = out.state[i]
        r = O.lib().oracle_kerr_radius(m, st[1], st[2], st[3])
        s = np.sqrt(r)
        ut = (r * s + a) / (r ** 0.75 * np.sqrt(r * s - 3 * s + 2 * a))
        g = obs_norm / (ut * (1 - Lz / E / (r * s + a)))
        worst_g = max(worst_g, abs(out.redshift[i] - g) / g)
        tau = O.lib().oracle_disk_temperature_ratio(r, a, r_in, 30.0)
        I = bins ** 3 / np.expm1(bins / (g * tau))
        worst_i = max(worst_i, np.abs(out.intensity[:, i] - I).max() / I.max())
    assert worst_g < 1e-6 and worst_i < 1e-5, (worst_g, worst_i)


def _formal_volumetric(m, o, x, p, obs_norm, bins, n=40001):
    """Independent formal solution of the volumetric transfer along one ray: the geodesic by an
    8th-order integrator with dense output, then I_k = int rho (nu_o^3 / nu_e^2) S(x_k / g)
    e^{-tau} dlam and tau = int kappa rho nu_e dlam by Simpson's rule on a uniform grid, with
    the emissivity definition of the trace options (rho = j0 r^-2 e^{-z^2 / (2 (h r)^2)} W(r),
    S(x) = x^3 / expm1(x / (6/r)^{3/4}), emitter on the Kepler circle r_c^2 = x^2 + y^2 - a^2, shell from the ISCO to r_max)."""
    from scipy.integrate import simpson
    a = m.a
    r_lo = max(o.vol_r_min, O.lib().oracle_isco(a))  # the emitting shell starts at the ISCO
    kr = lambda y: O.lib().oracle_kerr_radius(m, y[1], y[2], y[3])
    r_cap = 1.01 * O.lib().oracle_horizon(m)

    def rhs(lam, y):
        dx, dp, _ = O.rhs(m, y[:4], y[4:])
        return np.r_[dx, dp]
    cap = lambda lam, y: kr(y) - r_cap
    cap.terminal = True
    leave = lambda lam, y: kr(y) - 1.6 * o.vol_r_max
    leave.terminal, leave.direction = True, 1
    sol = solve_ivp(rhs, (0, 400), np.r_[x, p], method="DOP853", rtol=1e-11, atol=1e-13, dense_output=True,
                    events=[cap, leave], max_step=0.05)
    lam = np.linspace(0, sol.t[-1], n)
    Y = sol.sol(lam)
    ak, J = np.zeros(n), np.zeros((n, len(bins)))
    for j in range(n):
        y = Y[:, j]
        r = kr(y)
        if not (r_lo < r < o.vol_r_max):
            continue
        u = min(max((r - r_lo) / (0.1 * r_lo), 0), 1)
        v = min(max((o.vol_r_max - r) / (0.1 * o.vol_r_max), 0), 1)
        sig = o.vol_h_over_r * r
        window = u ** 3 * (u * (6 * u - 15) + 10) * v ** 3 * (v * (6 * v - 15) + 10)  # quintic smoothstep
        rho = o.vol_j0 / r ** 2 * np.exp(-y[3] ** 2 / (2 * sig ** 2)) * window
        rc = np.sqrt(max(y[1] ** 2 + y[2] ** 2 - a * a, 1e-12))
        Om = np.sqrt(rc) / (rc * rc + a * np.sqrt(rc))
        vel = np.array([1.0, -Om * y[2], Om * y[1], 0.0])
        vv = vel @ O.metric(m, y[:4])[0] @ vel
        if vv >= 0:
            continue
        nue = -(y[4] + Om * (y[1] * y[6] - y[2] * y[5])) / np.sqrt(-vv)
        nuo = -y[4] * obs_norm
        xe = bins * nue / nuo
        ak[j] = o.vol_kappa * rho * nue
        J[j] = rho * nuo ** 3 / nue ** 2 * xe ** 3 / np.expm1(xe / (6 / r) ** 0.75)
    tau = np.r_[0, np.cumsum(0.5 * (ak[1:] + ak[:-1]) * (lam[1] - lam[0]))]
    assert tau[-1] < o.vol_tau_max  # the rays used here never reach the opacity stop
    return simpson(J * np.exp(-tau)[:, None], x=lam, axis=0), tau[-1]


@pytest.mark.parametrize("kappa", [0.3, 30.0])
def test_volumetric_transfer_matches_formal_solution(kappa):
    """The trace's per-step 3-node Gauss-Legendre accumulation on the Hermite interpolant
    against the formal solution of the same emissivity along an independently integrated
    geodesic, optically thin (kappa 0.3, tau ~ 0.01) and thick (kappa 30, tau ~ 1)."""
    m = abi.metric_ks(1.0, 0.7)
    w, h = 12, 8
    cam = abi.camera(40.0, 65.0, 40.0, w, h)
    o = O.default_opts(abi.PRECISION_FP64)
    o.emission = abi.EMISSION_VOLUMETRIC
    o.n_bins = 8
    o.vol_kappa = kappa
    pix = np.array([1, 9, 60, 77])
    out, _ = O.trace([cam], [m], o, ray_ids=pix)
    bins = O.bin_centres(8, o.x_min, o.x_max)
    x_obs, _ = O.camera_state(cam, m, 0, 0)
    obs_norm = 1.0 / np.sqrt(-O.metric(m, x_obs)[0][0, 0])
    for j, i in enumerate(pix):
        assert out.cls[j] == abi.RAY_ESCAPED
        x, p = O.camera_state(cam, m, i % w, i // w)
        I, tau = _formal_volumetric(m, o, x, p, obs_norm, bins)
        assert abs(out.redshift[j] - tau) < 1e-5 * max(tau, 1e-3)  # optical depth travelled
        assert np.abs(out.intensity[:, j] - I).max() < 3e-5 * I.max(), (i, np.abs(out.intensity[:, j] - I).max() / I.max())
