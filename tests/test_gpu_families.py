"""GPU parity of the remaining spacetimes, integrators and emission models against the FP64
oracle, through the C-ABI.

Bars: the plain FP64 kernels equal the oracle's classes and agree to 1e-6 (rad for escape
directions, relative for intensities and optical depths); the FP32 path meets the trace
contract's bar (1e-4 rad, 1e-3 relative intensity against the pixel's brightest bin, floored
at 1e-2 of the median pixel) outside 0.5 M of the critical curve, where Kerr-Schild rays are
exempt.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from sirius_amd import abi
from workloads import random_spacetime_batch, warp_scene, wormhole_scene

pytestmark = pytest.mark.gpu


def _angle(u, v):
    c = (u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)
    return np.arccos(np.clip(c, -1, 1))


def _intensity_rel(got, ref, sel):
    Ir, Ig = ref.intensity[:, sel], got.intensity[:, sel]
    if Ir.size == 0:
        return 0.0
    scale = np.median(Ir.max(0))
    return float((np.abs(Ig - Ir).max(0) / np.maximum(Ir.max(0), 1e-2 * scale)).max())


def _check(got, ref, mask, ang_bar, rel_bar, flags=True):
    assert np.array_equal(got.cls[mask], ref.cls[mask])
    if flags:
        far = abi.FLAG_FAR_SIDE
        assert np.array_equal(got.flags[mask] & far, ref.flags[mask] & far)
    esc = mask & (ref.cls == abi.RAY_ESCAPED)
    ang = _angle(got.direction()[esc], ref.direction()[esc]).max() if esc.any() else 0.0
    rel = _intensity_rel(got, ref, mask & (ref.cls == abi.RAY_DISK))
    assert ang < ang_bar and rel < rel_bar, (ang, rel)
    return ang, rel


@pytest.mark.parametrize("scene", ["wormhole", "warp"])
def test_wormhole_and_warp_match_oracle(acc, scene):
    cam, m = (wormhole_scene if scene == "wormhole" else warp_scene)(96)
    o64 = O.default_opts(abi.PRECISION_FP64)
    ref, _ = O.trace([cam], [m], o64)
    everything = np.ones(ref.cls.shape, bool)
    print(f"[case] {scene} 96x96 fp64", flush=True)
    got, _ = acc.trace_batch([cam], [m], o64)
    _check(got, ref, everything, 1e-6, 1e-6)
    assert np.allclose(got.lam, ref.lam, rtol=1e-6)
    print(f"[case] {scene} 96x96 fp32", flush=True)
    got32, st = acc.trace_batch([cam], [m], O.default_opts(abi.PRECISION_FP32_ESCALATE))
    _check(got32, ref, everything, 1e-4, 1e-3)
    esc = (got32.flags & abi.FLAG_ESCALATED) != 0
    # warp: rays entering the bubble wall are re-traced in FP64; wormhole: rays whose impact
    # parameter is within 10 % of the throat; the rest stay FP32
    assert 0 < esc.sum() < 0.5 * esc.size and st["rays_escalated"] == esc.sum()
    if scene == "wormhole":
        far = (ref.flags & abi.FLAG_FAR_SIDE) != 0
        assert 0 < far.sum() < far.size  # both sides of the throat are seen


@pytest.mark.parametrize("scene", ["wormhole", "warp"])
def test_scene_frame_meets_the_bar_at_every_pixel(acc, scene):
    """The 1024x1024 wormhole and warp scenes with the library's default options: the FP32 path
    against the plain FP64 tracer at every pixel (classes, sides of the throat and escape
    directions within 1e-4 rad). The plain FP64 tracer's own agreement with the oracle is pinned
    at 96x96 above. The warp's wall crossers, re-traced in FP64, decide its maximum."""
    from sirius_amd.accelerator import default_opts
    cam, m = (wormhole_scene if scene == "wormhole" else warp_scene)(1024)
    o64, o32 = default_opts(abi.PRECISION_FP64), default_opts(abi.PRECISION_FP32_ESCALATE)
    o64.emission = o32.emission = abi.EMISSION_NONE
    print(f"[case] {scene} 1024x1024 fp64 and fp32 frames, every pixel", flush=True)
    ref, _ = acc.trace_batch([cam], [m], o64)
    got, _ = acc.trace_batch([cam], [m], o32)
    everything = np.ones(ref.cls.shape, bool)
    ang, _ = _check(got, ref, everything, 1e-4, 1e-3)
    print(f"[{scene} 1024] max escape-direction error {ang:.2e} rad", flush=True)


def test_symplectic_matches_oracle_and_reports_null_drift(acc):
    """Fixed-step Tao + Yoshida-4 (h = 0.01 M) on Kerr a = 0.9: the device and the oracle run
    the same fixed-step arithmetic, so classes, directions and null drift agree closely."""
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(30.0, 75.0, 30.0, 32, 18, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP64)
    o.integrator = abi.INTEGRATOR_SYMPLECTIC
    o.symplectic_h = 0.01
    o.r_escape = 60.0
    o.max_steps = 50000
    ref, _ = O.trace([cam], [m], o)
    print("[case] symplectic fp64 32x18", flush=True)
    got, st = acc.trace_batch([cam], [m], o)
    assert np.array_equal(got.cls, ref.cls)
    esc = ref.cls == abi.RAY_ESCAPED
    assert _angle(got.direction()[esc], ref.direction()[esc]).max() < 1e-6
    dk = ref.cls == abi.RAY_DISK
    assert np.allclose(got.intensity[:, dk], ref.intensity[:, dk], rtol=1e-6, atol=1e-12)
    assert np.array_equal(got.steps, ref.steps)
    # the null-constraint drift is reported per ray (disk rays at the Hermite-located crossing,
    # escaped rays at the escape step) and matches the oracle's ray by ray: both run the same
    # fixed-step arithmetic, and rounding moves a drift of ~1e-9 by a few per cent at most
    live = esc | dk
    gd, rd = got.null_drift[live], ref.null_drift[live]
    assert st["null_drift_max"] > 0 and np.median(gd) < 1e-8
    assert gd.max() < 1e-5 and rd.max() < 1e-5
    rel = np.abs(gd - rd) / np.maximum(rd, 1e-14)
    assert np.percentile(rel, 95) < 0.05, np.percentile(rel, 95)
    assert abs(np.median(gd) / np.median(rd) - 1) < 0.02 and abs(gd.max() / rd.max() - 1) < 0.05
    # the same rays by DP45: the symplectic result agrees to its step error
    od = O.default_opts(abi.PRECISION_FP64)
    od.r_escape = 60.0
    dp, _ = acc.trace_batch([cam], [m], od)
    both = esc & (dp.cls == abi.RAY_ESCAPED)
    assert np.percentile(_angle(got.direction()[both], dp.direction()[both]), 90) < 1e-4


def test_symplectic_1080p_kerr_frame_matches_oracle_on_a_pixel_sample(acc):
    """The 1920x1080 Kerr a = 0.9 camera (r_obs 50 M, inclination 75 deg, FOV 20 deg, disk from the
    ISCO to 30 M) by the fixed-step symplectic integrator, h = 0.01 M, FP64, escaping at 1000 M
    (about 1e5 steps per escaped ray). A fixed sample of pixels (a stride over the frame plus
    pixels in its top corners, where rays escape) is traced by the oracle: classes
    identical, escape directions, disk intensity and the per-ray null-constraint drift agree."""
    m = abi.metric_ks(1.0, 0.9)
    W, H = 1920, 1080
    cam = abi.camera(50.0, 75.0, 20.0, W, H, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP64)
    o.integrator = abi.INTEGRATOR_SYMPLECTIC
    o.symplectic_h = 0.01
    o.max_steps = 400000
    print("[case] symplectic fp64 1920x1080 h=0.01", flush=True)
    got, st = acc.trace_batch([cam], [m], o)
    assert (got.cls == abi.RAY_MAX_STEPS).sum() == 0 and (got.cls == abi.RAY_FAILED).sum() == 0
    # the frame's escaped rays lie in its top corners (1.5 % of the frame)
    corners = [y * W + x for y in (0, 20, 40) for x in (0, 40, 80, 120, W - 121, W - 81, W - 41, W - 1)]
    ids = np.unique(np.concatenate([np.arange(0, W * H, 20011), corners])).astype(np.int64)
    ref, _ = O.trace([cam], [m], o, ray_ids=ids)
    assert np.array_equal(got.cls[ids], ref.cls)
    esc = ref.cls == abi.RAY_ESCAPED
    dk = ref.cls == abi.RAY_DISK
    assert esc.sum() >= 10 and dk.sum() >= 50
    assert _angle(got.direction()[ids][esc], ref.direction()[esc]).max() < 1e-6
    assert np.allclose(got.intensity[:, ids][:, dk], ref.intensity[:, dk], rtol=1e-6, atol=1e-12)
    # disk rays (~5e3 steps): the drift is the integrator's truncation, the same on both sides;
    # escaped rays (~1e5 steps to 1000 M, where the metric is nearly flat): it is accumulated
    # rounding, which differs with every fused multiply-add, so both are only bounded
    gd, rd = got.null_drift[ids][dk], ref.null_drift[dk]
    rel = np.abs(gd - rd) / np.maximum(rd, 1e-14)
    assert np.percentile(rel, 90) < 0.05, np.percentile(rel, 90)
    assert abs(np.median(gd) / np.median(rd) - 1) < 0.05
    ge, re_ = got.null_drift[ids][esc], ref.null_drift[esc]
    print(f"[symplectic 1080p] escaped drift device max {ge.max():.2e}, oracle max {re_.max():.2e}", flush=True)
    assert ge.max() < 1e-8 and re_.max() < 1e-8
    # the frame's drift statistics: median over every ray, and the largest
    assert 0 < st["null_drift_median"] < 1e-8 and st["null_drift_max"] < 1e-2
    print(f"[symplectic 1080p] steps {st['steps_accepted']}, drift median {st['null_drift_median']:.2e}"
          f" max {st['null_drift_max']:.2e}", flush=True)


@pytest.mark.parametrize("prec", [abi.PRECISION_FP64, abi.PRECISION_FP32_ESCALATE])
def test_volumetric_emission_matches_oracle(acc, prec):
    m = abi.metric_ks(1.0, 0.7, 0.2, 5e-5)
    cam = abi.camera(40.0, 65.0, 30.0, 64, 48)
    o = O.default_opts(prec)
    o.emission = abi.EMISSION_VOLUMETRIC
    o.n_bins = 8
    ref, _ = O.trace([cam], [m], _fp64_like(o))
    print(f"[case] volumetric prec {prec} 64x48", flush=True)
    got, _ = acc.trace_batch([cam], [m], o)
    lit = ref.intensity.max(0) > 0
    assert lit.mean() > 0.3
    if prec == abi.PRECISION_FP64:
        assert np.array_equal(got.cls, ref.cls)
        assert np.allclose(got.intensity, ref.intensity, rtol=1e-6, atol=1e-9 * ref.intensity.max())
        assert np.allclose(got.redshift, ref.redshift, rtol=1e-6, atol=1e-9)  # optical depth
    else:
        crit = acc.critical_distance([cam], [m])
        keep = np.abs(crit) > 0.5
        assert np.array_equal(got.cls[keep], ref.cls[keep])
        sel = keep & lit & (got.cls == ref.cls)
        Ir, Ig = ref.intensity[:, sel], got.intensity[:, sel]
        rel = np.abs(Ig - Ir).max(0) / np.maximum(Ir.max(0), 1e-2 * np.median(Ir.max(0)))
        assert rel.max() < 1e-3, rel.max()


def _fp64_like(o):
    """FP64 oracle options with the emission settings of o."""
    r = O.default_opts(abi.PRECISION_FP64)
    for k in ("emission", "n_bins", "vol_j0", "vol_h_over_r", "vol_kappa", "vol_r_min", "vol_r_max", "vol_tau_max"):
        setattr(r, k, getattr(o, k))
    return r


def test_random_spacetime_batch_full_size(acc):
    """The 64-camera random Kerr-Newman-de Sitter batch (about 18.9 M rays, volumetric emission,
    8 bins) in one FP32 launch: every ray is classified, and a random sample of 3000 rays
    matches the FP64 oracle outside the critical band: classes identical, intensity within
    2e-3 (volumetric emission accumulates by per-step quadrature, and the FP32 path's steps are
    several times longer than the FP64 reference's; the thin-disk bar is 1e-3)."""
    cams, metrics = random_spacetime_batch()
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.emission = abi.EMISSION_VOLUMETRIC
    o.n_bins = 8
    n = sum(c.width * c.height for c in cams)
    print(f"[case] random batch 64 cameras {n} rays fp32", flush=True)
    st = acc.trace_resident(cams, metrics, o)
    assert st["rays"] == n
    got = acc.fetch_results()
    assert np.all(got.cls <= abi.RAY_DISK) and np.isfinite(got.intensity).all() and (got.intensity >= 0).all()
    rng = np.random.default_rng(0x51121005)
    ids = np.sort(rng.choice(n, 3000, replace=False))
    ref, _ = O.trace(cams, metrics, _fp64_like(o), ray_ids=ids)
    crit = acc.critical_distance(cams, metrics)[ids]
    sample = got.take(ids)
    keep = np.abs(crit) > 0.5
    assert np.array_equal(sample.cls[keep], ref.cls[keep])
    sel = keep & (ref.intensity.max(0) > 0)
    Ir, Ig = ref.intensity[:, sel], sample.intensity[:, sel]
    rel = np.abs(Ig - Ir).max(0) / np.maximum(Ir.max(0), 1e-2 * np.median(Ir.max(0)))
    assert rel.max() < 2e-3 and np.percentile(rel, 99) < 1e-3, (rel.max(), np.percentile(rel, 99))


def test_random_spacetime_batch_fp32_against_fp64_every_ray(acc):
    """The full 64-camera volumetric batch, every one of its 18.9 M rays: the FP32 path against
    the plain FP64 tracer outside the critical band. Classes identical; intensity relative to the
    pixel's brightest bin: 99 % within 1e-4 and 99.99 % within 2e-3, every ray within 5e-3. The
    tail beyond the sampled test's 2e-3 is about 450 of 17.6 M lit rays, all inside the shadow
    (critical distance -1.4 to -1.8 M: captured rays that cross the emitting layer on their way
    in); the FP32 path's step inside the layer (0.2 or 0.3 sigma) barely changes it (2.9e-3 and
    3.2e-3 at the worst ray). The FP64 tracer's own agreement with the oracle is pinned at 1/32 of
    the sides below."""
    from sirius_amd.accelerator import default_opts
    cams, metrics = random_spacetime_batch()
    o64, o32 = default_opts(abi.PRECISION_FP64), default_opts(abi.PRECISION_FP32_ESCALATE)
    for o in (o64, o32):
        o.emission = abi.EMISSION_VOLUMETRIC
        o.n_bins = 8
    print("[case] random batch 64 cameras fp64 and fp32, every ray", flush=True)
    acc.trace_resident(cams, metrics, o64, stats=False)
    ref = acc.fetch_results()
    acc.trace_resident(cams, metrics, o32, stats=False)
    got = acc.fetch_results()
    keep = np.abs(acc.critical_distance(cams, metrics)) > 0.5
    assert np.array_equal(got.cls[keep], ref.cls[keep])
    sel = keep & (ref.intensity.max(0) > 0)
    Ir, Ig = ref.intensity[:, sel], got.intensity[:, sel]
    rel = np.abs(Ig - Ir).max(0) / np.maximum(Ir.max(0), 1e-2 * np.median(Ir.max(0)))
    print(f"[batch every ray] {sel.sum()} lit rays, intensity error max {rel.max():.2e}, "
          f"p99 {np.percentile(rel, 99):.2e}, p99.99 {np.percentile(rel, 99.99):.2e}, "
          f"above 2e-3 {(rel > 2e-3).sum()}", flush=True)
    assert np.percentile(rel, 99) < 1e-4 and np.percentile(rel, 99.99) < 2e-3 and rel.max() < 5e-3, \
        (np.percentile(rel, 99), np.percentile(rel, 99.99), rel.max())


def test_random_spacetime_batch_fp64_matches_oracle(acc):
    """The same batch at 1/32 of the image sides, every ray, plain FP64 against the oracle."""
    cams, metrics = random_spacetime_batch(scale=32)
    o = O.default_opts(abi.PRECISION_FP64)
    o.emission = abi.EMISSION_VOLUMETRIC
    o.n_bins = 8
    ref, _ = O.trace(cams, metrics, o)
    print(f"[case] random batch 64 cameras {ref.cls.size} rays fp64", flush=True)
    got, _ = acc.trace_batch(cams, metrics, o)
    assert np.array_equal(got.cls, ref.cls)
    assert np.allclose(got.intensity, ref.intensity, rtol=1e-6, atol=1e-9 * ref.intensity.max())


def test_autotuner_picks_a_launch_shape_that_keeps_results(acc):
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, 480, 270, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.escalate_margin = 0.0
    base, _ = acc.trace_batch([cam], [m], o)
    print("[case] autotune 480x270", flush=True)
    best = acc.autotune([cam], [m], o)
    assert best > 0
    assert o.block_size in (64, 128, 256) and o.blocks_per_cu >= 1 and 1 <= o.refill_min <= 64
    tuned, _ = acc.trace_batch([cam], [m], o)
    for k in ("cls", "theta", "phi", "steps", "intensity"):
        assert np.array_equal(getattr(tuned, k), getattr(base, k)), k
