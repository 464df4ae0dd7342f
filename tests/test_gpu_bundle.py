"""Ray bundles on the device, through the C-ABI.

1. The bundle's Jacobian of the escape direction d(theta, phi)/d(u, v) equals the exact
   Schwarzschild lens map's (quadrature of the orbit equation, no integrator), and in Kerr
   (a = 0.9, with the thin disk: escape directions and disk hit points), the Alcubierre bubble and
   the Morris-Thorne wormhole the oracle's central differences of whole FP64 traces at sub-pixel
   offsets; rays within 1 M of the critical curve, where the map is not smooth on a pixel, are
   left out.
2. With the bundle enabled, launch's star light per pixel is the average over the pixel of the
   point-sampled light (checked against 8 x 8 supersampling) for stars whose spread covers a few
   pixels; without it, launch point-samples as before.
"""
import numpy as np
import pytest

from exact_lensing import schwarzschild_jacobian
from oracle import oracle_py as O
from sirius_amd import abi
from sirius_amd.accelerator import starfield_default

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-30)


def test_bundle_matches_exact_schwarzschild(acc):
    m = abi.metric_ks(1.0, 0.0)
    cam = abi.camera(30.0, 80.0, 40.0, 24, 16)
    o = O.default_opts(abi.PRECISION_FP64)
    o.emission, o.n_bins = abi.EMISSION_NONE, 0
    print("[case] bundle schwarzschild 24x16", flush=True)
    b = acc.trace_bundle([cam], [m], o)
    errs = []
    for pid in range(0, 24 * 16, 3):
        px, py = pid % 24, pid // 24
        if b["cls"][pid] != abi.RAY_ESCAPED or abs(O.critical_distance(cam, m, px, py)) < 1.0:
            continue
        Je = schwarzschild_jacobian(cam, m, px, py)
        if Je is None:
            continue
        errs.append(np.linalg.norm(b["jac"][:, pid] - Je) / np.linalg.norm(Je))
    assert len(errs) > 40
    assert np.median(errs) < 1e-4 and max(errs) < 2e-3, (np.median(errs), max(errs))


@pytest.mark.parametrize("scene", ["kerr_disk", "warp", "wormhole"])
def test_bundle_matches_oracle_differences(acc, scene):
    o = O.default_opts(abi.PRECISION_FP64)
    if scene == "kerr_disk":
        m = abi.metric_ks(1.0, 0.9)
        cam = abi.camera(50.0, 75.0, 30.0, 32, 18, disk=(0.0, 30.0))
        o.emission, o.n_bins = abi.EMISSION_THIN_DISK, 0
    elif scene == "warp":
        from workloads import warp_scene
        cam, m = warp_scene(24)
        o.emission, o.n_bins = abi.EMISSION_NONE, 0
    else:
        from workloads import wormhole_scene
        cam, m = wormhole_scene(24)
        o.emission, o.n_bins = abi.EMISSION_NONE, 0
    print(f"[case] bundle vs oracle differences: {scene}", flush=True)
    b = acc.trace_bundle([cam], [m], o)
    n = cam.width * cam.height
    ids = np.arange(n, dtype=np.int64)
    ref_o = O.default_opts(abi.PRECISION_FP64)
    ref_o.rtol, ref_o.atol, ref_o.emission, ref_o.n_bins = 1e-11, 1e-13, o.emission, 0
    if scene == "wormhole":
        # the wormhole's sky direction needs its orbital-plane map: differences of the oracle's
        # escape angles (float) over a wider offset
        eps = 0.05
        base = O.trace([cam], [m], ref_o, ray_ids=ids)[0]
        J = np.zeros((4, n))
        same = np.ones(n, bool)
        for a in (0, 1):
            p = O.trace_subpixel([cam], [m], ref_o, ids, eps if a == 0 else 0, eps if a == 1 else 0)
            q = O.trace_subpixel([cam], [m], ref_o, ids, -eps if a == 0 else 0, -eps if a == 1 else 0)
            same &= (p.cls == base.cls) & (q.cls == base.cls) & ((p.flags & 16) == (q.flags & 16))
            J[a] = (p.theta.astype(float) - q.theta) / (2 * eps)
            J[2 + a] = (np.remainder(p.phi.astype(float) - q.phi + np.pi, 2 * np.pi) - np.pi) / (2 * eps)
        cls = base.cls
        bar_med, bar_90 = 2e-3, 2e-2
    else:
        J, cls, same = O.lens_jacobian([cam], [m], ref_o, ids)
        bar_med, bar_90 = 1e-4, 2e-3
    keep = same & (cls == b["cls"]) & ((cls == abi.RAY_ESCAPED) | (cls == abi.RAY_DISK))
    if scene == "kerr_disk":
        crit = np.array([O.critical_distance(cam, m, int(i % cam.width), int(i // cam.width)) for i in ids])
        keep &= np.abs(crit) > 1.0
    assert keep.sum() > 0.3 * n, keep.sum()
    err = _rel(b["jac"][:, keep].astype(float), J[:, keep])
    assert np.median(err) < bar_med and np.percentile(err, 90) < bar_90, (np.median(err), np.percentile(err, 90))
    if scene == "kerr_disk":
        assert (b["cls"][keep] == abi.RAY_DISK).sum() > 20 and (b["cls"][keep] == abi.RAY_ESCAPED).sum() > 20
    # beams: semi-axes ordered, finite; captured rays terminated
    assert np.all(b["beam"][0] >= b["beam"][1]) and np.isfinite(b["beam"]).all()
    assert np.all(b["flags"][b["cls"] == abi.RAY_CAPTURED] & abi.BUNDLE_TERMINATED)


def _blur_catalogue(rng, n=6000):
    """Stars whose spread covers one to three pixels of a 20-degree, 54-pixel-high view (6.5e-3 rad
    a pixel): distances of 2.5 to 6 pc with a 50 mm aperture focused at 100 pc
    (sigma = 0.05 |1/d - 1/100| rad); about a hundred fall in the view."""
    rows = np.zeros((n, 8), np.float32)
    v = rng.normal(size=(n, 3))
    rows[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    rows[:, 3] = rng.uniform(2.5, 6.0, n)
    rows[:, 4] = rng.uniform(-1.0, 3.0, n)
    rows[:, 5] = 0.6
    rows[:, 6] = rng.uniform(3000.0, 12000.0, n)
    arr = (abi.Star * n)()
    np.frombuffer(arr, np.float32).reshape(-1, 8)[:] = rows
    return arr


def test_beam_star_light_is_the_pixel_average(acc):
    m = abi.metric_ks(1.0, 0.9)
    w, h, k = 96, 54, 8
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.emission, o.n_bins = abi.EMISSION_NONE, 0
    cfg = starfield_default()
    cfg.aperture_mm, cfg.focus_distance_pc, cfg.enable_dof, cfg.brightness_scale = 50.0, 100.0, 1, 1.0
    stars = _blur_catalogue(np.random.default_rng(3))
    acc.set_starfield(0, cfg, stars, len(stars))
    acc.set_starfield(1, None, None, 0)
    try:
        print("[case] supersampled 8x8 frame", flush=True)
        big = abi.camera(50.0, 75.0, 20.0, w * k, h * k)
        acc.initialise(w * k, h * k)
        acc.uploadBackground(np.zeros((1, 1, 4), np.uint8))
        acc.set_ray_bundle(False)
        acc.launch(big, m, o)
        ss = acc.getFrameBuffer().reshape(h, k, w, k, 4)[..., :3]
        rays, _ = acc.trace_batch([big], [m], o)
        all_esc = (rays.cls.reshape(h, k, w, k) == abi.RAY_ESCAPED).all(axis=(1, 3))
        ss = ss.mean(axis=(1, 3))
        small = abi.camera(50.0, 75.0, 20.0, w, h)
        acc.initialise(w, h)
        acc.uploadBackground(np.zeros((1, 1, 4), np.uint8))
        print("[case] point-sampled and beam frames", flush=True)
        acc.launch(small, m, o)
        point = acc.getFrameBuffer().reshape(h, w, 4)[..., :3].copy()
        acc.set_ray_bundle(True)
        acc.launch(small, m, o)
        beam = acc.getFrameBuffer().reshape(h, w, 4)[..., :3].copy()
        acc.set_ray_bundle(False)
        acc.launch(small, m, o)
        again = acc.getFrameBuffer().reshape(h, w, 4)[..., :3]
        assert np.array_equal(again, point)  # disabling restores point sampling exactly
    finally:
        acc.set_starfield(0, None, None, 0)
        acc.set_ray_bundle(False)
    # The beam is a Gaussian footprint of the pixel under the lens map's linearisation: exact where
    # the map is close to linear across the pixel (6 M or more from the critical curve here: a
    # pixel spans 6.5e-3 rad), and still far nearer the pixel average than a centre sample where
    # the map bends within the pixel (3 to 6 M).
    crit = np.array([O.critical_distance(small, m, i % w, i // w) for i in range(w * h)]).reshape(h, w)
    lit = all_esc & (ss.max(-1) > 1e-3 * ss.max())
    e_b = np.abs(beam - ss).max(-1) / np.maximum(ss.max(-1), 1e-30)
    e_p = np.abs(point - ss).max(-1) / np.maximum(ss.max(-1), 1e-30)
    far = lit & (np.abs(crit) >= 6.0)
    mid = lit & (np.abs(crit) >= 3.0) & (np.abs(crit) < 6.0)
    print(f"far: {far.sum()} px, beam median {np.median(e_b[far]):.4f} p90 {np.percentile(e_b[far], 90):.4f}, "
          f"point median {np.median(e_p[far]):.4f}; 3-6 M: {mid.sum()} px, beam median "
          f"{np.median(e_b[mid]):.4f}, point median {np.median(e_p[mid]):.4f}", flush=True)
    assert far.sum() > 500 and mid.sum() > 300
    assert np.median(e_b[far]) < 0.005 and np.percentile(e_b[far], 90) < 0.03
    assert np.median(e_b[far]) < 0.1 * np.median(e_p[far])
    assert np.median(e_b[mid]) < 0.2 * np.median(e_p[mid])
    tot = lambda f, sel: f[sel].sum()
    assert abs(tot(beam, far) - tot(ss, far)) < 0.01 * tot(ss, far)


def test_bundle_batch_of_cameras_equals_single_cameras(acc):
    """A batch of two cameras (different spins, sizes and metrics) gives each camera's bundles as
    tracing it alone does, bit for bit."""
    m0, m1 = abi.metric_ks(1.0, 0.5), abi.metric_ks(1.0, 0.95, 0.2)
    c0 = abi.camera(40.0, 60.0, 30.0, 20, 12, disk=(0.0, 20.0), metric_index=0)
    c1 = abi.camera(25.0, 85.0, 35.0, 16, 16, disk=(0.0, 20.0), metric_index=1)
    o = O.default_opts(abi.PRECISION_FP64)
    o.emission, o.n_bins = abi.EMISSION_THIN_DISK, 0
    print("[case] bundle batch of two cameras", flush=True)
    both = acc.trace_bundle([c0, c1], [m0, m1], o)
    c1s = abi.camera(25.0, 85.0, 35.0, 16, 16, disk=(0.0, 20.0), metric_index=0)
    one = acc.trace_bundle([c0], [m0], o)
    two = acc.trace_bundle([c1s], [m1], o)
    n0 = 20 * 12
    for k in ("cls", "flags"):
        assert np.array_equal(both[k][:n0], one[k]) and np.array_equal(both[k][n0:], two[k]), k
    for k in ("jac", "beam"):
        assert np.array_equal(both[k][:, :n0], one[k]) and np.array_equal(both[k][:, n0:], two[k]), k
