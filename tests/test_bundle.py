"""Ray bundles, on the CPU: the oracle's derivative of traced outcomes with respect to the image
point (central differences of whole FP64 traces at sub-pixel offsets), pinned by

- the offset plumbing: offset (0, 0) is the pixel-centre trace and offset (1, 0) is the next
  pixel's trace, bit for bit;
- the exact Schwarzschild lens map: escape directions by quadrature of the orbit equation, with
  no integrator, differentiated the same way (tests/exact_lensing.py);
- the symmetry of Schwarzschild lensing: rotating the image point about the image centre (the
  camera looks at the hole) rotates the escape direction about the hole-observer axis by the same
  angle, so J applied to the rotation generator is that rotation's generator at the escape
  direction.
"""
import numpy as np

from exact_lensing import schwarzschild_jacobian
from oracle import oracle_py as O
from sirius_amd import abi


def _opts():
    o = O.default_opts(abi.PRECISION_FP64)
    o.rtol, o.atol, o.emission, o.n_bins = 1e-11, 1e-13, abi.EMISSION_NONE, 0
    return o


def test_subpixel_offset_plumbing():
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(40.0, 70.0, 25.0, 16, 9)
    o = O.default_opts(abi.PRECISION_FP64)
    ids = np.array([3, 20, 50, 77, 130], np.int64)
    base = O.trace([cam], [m], o, ray_ids=ids, with_state=True)[0]
    same = O.trace_subpixel([cam], [m], o, ids, 0.0, 0.0)
    assert np.array_equal(base.state, same.state) and np.array_equal(base.cls, same.cls)
    shifted = O.trace_subpixel([cam], [m], o, ids, 1.0, 0.0)
    nxt = O.trace([cam], [m], o, ray_ids=ids + 1, with_state=True)[0]
    assert np.array_equal(shifted.state, nxt.state)
    # and the offset is cleared afterwards
    again = O.trace([cam], [m], o, ray_ids=ids, with_state=True)[0]
    assert np.array_equal(again.state, base.state)


def _schwarzschild_case():
    m = abi.metric_ks(1.0, 0.0)
    cam = abi.camera(30.0, 80.0, 40.0, 24, 16)
    rng = np.random.default_rng(7)
    ids = np.sort(rng.choice(24 * 16, 40, replace=False)).astype(np.int64)
    return m, cam, ids


def test_oracle_jacobian_matches_exact_schwarzschild():
    m, cam, ids = _schwarzschild_case()
    J, cls, same = O.lens_jacobian([cam], [m], _opts(), ids)
    n_ok = 0
    for i, pid in enumerate(ids):
        if cls[i] != abi.RAY_ESCAPED or not same[i]:
            continue
        if abs(O.critical_distance(cam, m, int(pid % cam.width), int(pid // cam.width))) < 1.0:
            continue
        Je = schwarzschild_jacobian(cam, m, int(pid % cam.width), int(pid // cam.width))
        if Je is None:
            continue
        assert np.linalg.norm(J[:, i] - Je) < 1e-5 * np.linalg.norm(Je), (pid, J[:, i], Je)
        n_ok += 1
    assert n_ok >= 15


def test_schwarzschild_jacobian_is_equivariant_under_rotation():
    m, cam, ids = _schwarzschild_case()
    J, cls, same = O.lens_jacobian([cam], [m], _opts(), ids)
    base = O.trace([cam], [m], _opts(), ray_ids=ids, with_state=True)[0]
    ang, _ = O.outcome_coords(m, base)
    er = np.array([np.sin(cam.theta_obs) * np.cos(cam.phi_obs), np.sin(cam.theta_obs) * np.sin(cam.phi_obs),
                   np.cos(cam.theta_obs)])
    n_ok = 0
    for i, pid in enumerate(ids):
        if cls[i] != abi.RAY_ESCAPED or not same[i]:
            continue
        u, v = pid % cam.width + 0.5, pid // cam.width + 0.5
        w = np.array([-(v - cam.height / 2), u - cam.width / 2])  # image rotation generator (pixels)
        th, ph = ang[0, i], ang[1, i]
        e_th = np.array([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)])
        e_ph = np.array([-np.sin(ph), np.cos(ph), 0.0])
        d = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        got = (J[0, i] * w[0] + J[1, i] * w[1]) * e_th + np.sin(th) * (J[2, i] * w[0] + J[3, i] * w[1]) * e_ph
        gen = np.cross(er, d)
        # the image rotation maps to a rotation about +-er of the same angle: got = +-gen
        err = min(np.linalg.norm(got - gen), np.linalg.norm(got + gen))
        assert err < 5e-4 * max(np.linalg.norm(gen), 1e-3), (pid, got, gen)  # difference truncation
        n_ok += 1
    assert n_ok >= 15
