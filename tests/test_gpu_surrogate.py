"""The far-field surrogate in the FP32 path, through the C-ABI: network inputs, training on
integrator labels, the accept / fall-back decision, the audit, and the ablation (the same
frame with the surrogate off).

What is exact: rays the surrogate does not answer are traced by the same FP32 kernel as with
the surrogate off, so their results are bit-identical; the audit re-traces answered rays with
the escalation policy and its reported error equals the error measured here. What is
statistical: the network's accuracy, which the audit measures and the decision thresholds
(fallback band, class probability) bound.
"""
import numpy as np
import pytest

from devmem import DeviceArray
from oracle import oracle_py as O
from sirius_amd import abi
from workloads import MT19937_64

pytestmark = pytest.mark.gpu


def _angle(u, v):
    c = (u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)
    return np.arccos(np.clip(c, -1, 1))


def test_features_match_oracle_camera(acc):
    m = abi.metric_ks(1.0, 0.7)
    cam = abi.camera(40.0, 60.0, 20.0, 24, 16)
    xd = DeviceArray(shape=(24 * 16, 5), dtype=np.float32)
    nd = DeviceArray(shape=(24 * 16, 3), dtype=np.float32)
    acc.surrogate_features([cam], [m], xd.data_ptr(), nd.data_ptr())
    x, n0 = xd.get(), nd.get()
    for i in range(24 * 16):
        xs, ps = O.camera_state(cam, m, i % 24, i // 24)
        _, gi, _ = O.metric(m, xs)
        k = gi @ ps
        assert np.allclose(n0[i], k[1:] / np.linalg.norm(k[1:]), atol=1e-6)
        # Bardeen impact parameters from the conserved E, L_z and the polar momentum
        E, L = -ps[0], xs[1] * ps[2] - xs[2] * ps[1]
        r = O.lib().oracle_kerr_radius(m, xs[1], xs[2], xs[3])
        ct = xs[3] / r
        st = np.sqrt(1 - ct * ct)
        pth = (ct / st) * (xs[1] * ps[1] + xs[2] * ps[2]) - r * st * ps[3]
        want = [0.7, np.deg2rad(60.0) / (np.pi / 2), (40.0 - 60.0) / 40.0, -L / (E * st) / 15.0, pth / E / 15.0]
        assert np.allclose(x[i], want, atol=2e-6), (i, x[i], want)


def _training_cameras(n_cam, side, seed):
    g = MT19937_64(seed)
    cams, metrics = [], []
    for i in range(n_cam):
        a = g.uniform(0.0, 0.99)
        incl = g.uniform(5.0, 89.0)
        r_obs = g.uniform(20.0, 100.0)
        metrics.append(abi.metric_ks(1.0, a))
        cams.append(abi.camera(r_obs, incl, 40.0, side, side, disk=(0.0, 30.0), metric_index=i))
    return cams, metrics


def test_surrogate_training_decision_audit_and_ablation(acc):
    cams, metrics = _training_cameras(64, 64, 0x51121006)
    lab = O.default_opts(abi.PRECISION_FP64)
    lab.far_slope = 0.25
    acc.mlp_init(0x51121006)
    print("[case] surrogate train 64 cameras x 64^2", flush=True)
    first = acc.surrogate_train(cams, metrics, lab, 1, 4096, 2e-3, 1)
    last = acc.surrogate_train(cams, metrics, lab, 20, 4096, 2e-3, 2)
    print(f"[surrogate] loss {first:.4f} -> {last:.4f}", flush=True)
    assert np.isfinite(last) and last < 0.5 * first, (first, last)

    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(60.0, 60.0, 40.0, 160, 90, disk=(0.0, 30.0))
    off = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    print("[case] surrogate off 160x90", flush=True)
    base, st_off = acc.trace_batch([cam], [m], off)
    on = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    on.use_surrogate = 1
    on.surrogate_pmin = 0.9
    on.audit_fraction = 1.0
    print("[case] surrogate on 160x90", flush=True)
    got, st = acc.trace_batch([cam], [m], on)
    sur = (got.flags & abi.FLAG_SURROGATE) != 0
    assert not ((base.flags & abi.FLAG_SURROGATE) != 0).any()
    assert st["rays_surrogate"] == sur.sum() and st["rays_fallback"] == (~sur).sum()
    assert sur.sum() > 0 and (~sur).sum() > 0
    # fallback rays: the same FP32 arithmetic as with the surrogate off, bit for bit
    for k in ("cls", "theta", "phi", "lam", "disk_r", "steps", "intensity"):
        assert np.array_equal(getattr(got, k)[..., ~sur], getattr(base, k)[..., ~sur]), k
    # answered rays: outside the fallback band, never the disk class, no integration
    crit = acc.critical_distance([cam], [m])
    assert np.all(np.abs(crit[sur]) >= on.surrogate_band - 1e-4)
    assert np.all(got.cls[sur] != abi.RAY_DISK) and np.all(got.steps[sur] == 0)
    # the audit (every answer, audit_fraction 1) reports the error measured against the
    # integrator's own answer for the same rays
    assert st["rays_audited"] == sur.sum()
    o64 = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o64.precision = abi.PRECISION_FP64
    ertol = on.escalate_rtol or on.rtol  # 0: escalations keep the path's tolerance
    o64.rtol, o64.atol = ertol, on.atol * (ertol / on.rtol)
    ref, _ = acc.trace_batch([cam], [m], o64)
    same = sur & (ref.cls == got.cls)
    assert st["audit_class_mismatch"] == (sur & (ref.cls != got.cls)).sum()
    esc = same & (got.cls == abi.RAY_ESCAPED)
    err = _angle(got.direction()[esc], ref.direction()[esc])
    assert abs(st["audit_dir_err_max"] - err.max()) < 1e-5 * max(1.0, err.max())
    print(f"[surrogate] answered {sur.mean():.3f}, class agreement {(ref.cls[sur] == got.cls[sur]).mean():.4f}, "
          f"direction error median {np.median(err):.2e} max {err.max():.2e} rad", flush=True)
    assert (ref.cls[sur] == got.cls[sur]).mean() > 0.95
    assert np.median(err) < 5e-2
