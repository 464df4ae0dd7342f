"""Pins of the oracle's metrics (oracle/sirius_oracle.c) against the reference's own metric
code (golden vectors from KerrSchildFamily/MorrisThorneFamily/WarpDriveFamily::evaluate,
PHMT100A.h:248-396, PHMT101A.h:285-345, PHMT102A.h:515-578), against central finite
differences (the reference's own check, TSPH012A.cpp:78-90) and against the closed-form
Kerr-Schild inverse g^{mu nu} = eta - H l^mu l^nu."""
import numpy as np

from oracle import oracle_py as O
from sirius_amd import abi


def test_kerr_schild_matches_reference(golden):
    d = golden("ref_metric.npz")
    for (M, a, Q, L), x, g_ref, dg_ref in zip(d["ks_params"], d["ks_x"], d["ks_g"], d["ks_dg"]):
        g, gi, dg = O.metric(abi.metric_ks(M, a, Q, L), x)
        assert np.allclose(g, g_ref, rtol=0, atol=1e-13 * max(1, abs(g_ref).max()))
        assert np.allclose(dg, dg_ref, rtol=0, atol=1e-12 * max(1, abs(dg_ref).max()))


def test_kerr_schild_gradient_finite_difference():
    rng = np.random.default_rng(7)
    for _ in range(50):
        m = abi.metric_ks(1.0, rng.uniform(-0.95, 0.95), rng.uniform(0, 0.3), rng.choice([0, 5e-4]))
        x = np.r_[0.0, rng.uniform(-15, 15, 3)]
        if np.linalg.norm(x[1:]) < 3:
            continue
        _, _, dg = O.metric(m, x)
        for lam in (1, 2, 3):
            e = np.zeros(4); e[lam] = 1e-5
            fd = (O.metric(m, x + e)[0] - O.metric(m, x - e)[0]) / 2e-5
            assert np.allclose(dg[lam], fd, atol=1e-7)
        assert np.all(dg[0] == 0)


def test_kerr_schild_closed_form_inverse():
    """g^{-1} = eta - H l l with l^mu = (-1, l_1, l_2, l_3): checks the oracle's numerical
    inverse and the metric's Kerr-Schild structure at once (Lambda = 0)."""
    rng = np.random.default_rng(8)
    for _ in range(40):
        M, a, Q = 1.0, rng.uniform(-0.99, 0.99), rng.uniform(0, 0.3)
        m = abi.metric_ks(M, a, Q)
        x = np.r_[0.0, rng.uniform(-12, 12, 3)]
        if np.linalg.norm(x[1:]) < 3:
            continue
        X, Y, Z = x[1:]
        R2 = X * X + Y * Y + Z * Z
        r = np.sqrt(0.5 * (R2 - a * a + np.sqrt((R2 - a * a) ** 2 + 4 * a * a * Z * Z)))
        H = (2 * M * r - Q * Q) * r * r / (r ** 4 + a * a * Z * Z)
        l_lo = np.array([1, (r * X + a * Y) / (r * r + a * a), (r * Y - a * X) / (r * r + a * a), Z / r])
        eta = np.diag([-1.0, 1, 1, 1])
        l_up = eta @ l_lo
        g, gi, _ = O.metric(m, x)
        assert np.allclose(g, eta + H * np.outer(l_lo, l_lo), atol=1e-13)
        assert np.allclose(gi, eta - H * np.outer(l_up, l_up), atol=1e-12)
        assert abs(l_up @ l_lo) < 1e-12          # l is null
        assert abs(np.linalg.det(g) + 1) < 1e-10  # det g = -1 for Kerr-Schild


def test_kerr_radius_and_horizon():
    for a in (0.0, 0.5, 0.9, 0.999):
        m = abi.metric_ks(1.0, a)
        rp = O.lib().oracle_horizon(m)
        assert abs(rp - (1 + np.sqrt(1 - a * a))) < 1e-15
        # equatorial Kerr-Schild point at cylindrical radius rho has r = sqrt(rho^2 - a^2)
        assert abs(O.lib().oracle_kerr_radius(m, 5.0, 0.0, 0.0) - np.sqrt(25 - a * a)) < 1e-12
        # on the axis r = |z|
        assert abs(O.lib().oracle_kerr_radius(m, 0.0, 0.0, 7.0) - 7.0) < 1e-12
    assert abs(O.lib().oracle_horizon(abi.metric_ks(1.0, 0.999)) - 1.0447) < 1e-4  # TSVAL001A


def test_wormhole_matches_reference_in_proper_length(golden):
    """The oracle's Ellis wormhole is written in l with r^2 = l^2 + b0^2, reduced to the orbital
    plane; the reference's is in r (PHMT101A.h). Same spacetime: g_ll = 1 <=> g_rr dr^2 = dl^2,
    g_psipsi(l) = g_thth(r) = r^2, d g_psipsi/dl = (d g_thth/dr)(dr/dl)."""
    d = golden("ref_metric.npz")
    b0 = float(d["mt_b0"])
    m = abi.metric_mt(b0)
    for x, g_ref, dg_ref in zip(d["mt_x"], d["mt_g"], d["mt_dg"]):
        r = max(x[1], 1.001 * b0)
        l = np.sqrt(r * r - b0 * b0)
        g, gi, dg = O.metric(m, np.array([0.0, l, 0.0, 0.0]))
        assert abs(g[0, 0] - g_ref[0, 0]) < 1e-14
        assert abs(g_ref[1, 1] - r * r / (r * r - b0 * b0)) < 1e-9 * g_ref[1, 1]
        assert abs(g[2, 2] - g_ref[2, 2]) < 1e-12 * g_ref[2, 2]
        assert abs(dg[1, 2, 2] - dg_ref[1, 2, 2] * (l / r)) < 1e-12 * abs(dg_ref[1, 2, 2])
        assert np.allclose(gi @ g, np.eye(4), atol=1e-14)


def test_alcubierre_matches_reference(golden):
    d = golden("ref_metric.npz")
    v, R, s = d["warp_params"]
    m = abi.metric_warp(v, R, s)
    for x, g_ref, dg_ref in zip(d["warp_x"], d["warp_g"], d["warp_dg"]):
        g, gi, dg = O.metric(m, x)
        assert np.allclose(g, g_ref, atol=1e-13)
        assert np.allclose(dg, dg_ref, atol=1e-11 * max(1, abs(dg_ref).max()))
        assert np.allclose(gi @ g, np.eye(4), atol=1e-12)


def test_rhs_is_hamiltons_equations():
    """dx = dH/dp and dp = -dH/dx for H = 1/2 g^{mu nu} p p, checked by finite differences of
    H itself (independent of the oracle's closed forms)."""
    rng = np.random.default_rng(9)
    m = abi.metric_ks(1.0, 0.7, 0.2, 1e-4)

    def Ham(x, p):
        return 0.5 * p @ O.metric(m, x)[1] @ p
    for _ in range(20):
        x = np.r_[0.0, rng.uniform(-10, 10, 3)]
        if np.linalg.norm(x[1:]) < 3:
            continue
        p = rng.normal(size=4)
        dx, dp, pk = O.rhs(m, x, p)
        for i in range(4):
            e = np.zeros(4); e[i] = 1e-6
            assert abs(dx[i] - (Ham(x, p + e) - Ham(x, p - e)) / 2e-6) < 1e-6 * max(1, abs(dx[i]))
            if i > 0:
                assert abs(dp[i] + (Ham(x + e, p) - Ham(x - e, p)) / 2e-6) < 1e-6 * max(1, abs(dp[i]))
        assert abs(pk - 2 * Ham(x, p)) < 1e-12 * max(1, abs(pk))
