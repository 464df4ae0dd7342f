"""GPU parity of the HIP tracer (through the C-ABI) against the FP64 oracle and its committed
golden columns.

Bars (the trace contract's tolerances):
  plain FP64 kernel   classes identical; escape direction within 1e-6 rad; disk radius and
                      intensity within 1e-6 relative (same algorithm and arithmetic width as
                      the oracle; differences are FMA contraction and libm ulps)
  FP32 + escalation   classes identical except rays within 0.5 M of the critical curve;
                      escape direction within 1e-4 rad; intensity within 1e-3 relative to the
                      pixel's brightest bin, floored at 1e-2 of the frame's median disk pixel
                      (intensity vanishes at the ISCO, where a relative error means nothing)
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from sirius_amd import abi
from sirius_amd.accelerator import SiriusHipError

pytestmark = pytest.mark.gpu


def _angle(u, v):
    u = u / np.linalg.norm(u, axis=1, keepdims=True)
    v = v / np.linalg.norm(v, axis=1, keepdims=True)
    return np.arccos(np.clip((u * v).sum(1), -1, 1))


def _dirs(th, ph):
    th, ph = np.asarray(th, float), np.asarray(ph, float)
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)


def _setup(d):
    w, h = int(d["width"]), int(d["height"])
    cam = abi.camera(float(d["r_obs"]), float(d["incl_deg"]), float(d["fov"]), w, h, disk=tuple(d["disk"]))
    m = abi.metric_ks(1.0, float(d["a"]))
    return cam, m


def _compare(got, d, band, fp64):
    """got: HostRays restricted to the fixture pixels; d: golden fixture."""
    cls_ref = d["fp64_cls"]
    if fp64:
        assert np.array_equal(got.cls, cls_ref)
        mask = np.ones_like(band)
    else:
        assert np.array_equal(got.cls[band], cls_ref[band])
        mask = band
    esc = mask & (got.cls == 0) & (cls_ref == 0)
    ang = _angle(got.direction()[esc], _dirs(d["fp64_theta"][esc], d["fp64_phi"][esc])) if esc.any() else np.zeros(1)
    dk = mask & (got.cls == 2) & (cls_ref == 2)
    Iref = d["fp64_I"][:, dk]
    Igot = got.intensity[:, dk]
    scale = np.median(Iref.max(0))
    rel = np.abs(Igot - Iref).max(0) / np.maximum(Iref.max(0), 1e-2 * scale)
    rdisk = np.abs(got.disk_r[dk] - d["fp64_disk_r"][dk]) / d["fp64_disk_r"][dk]
    return ang.max(), rel.max(), rdisk.max() if rdisk.size else 0.0


@pytest.mark.parametrize("fixture", ["ref_trace_c1.npz", "ref_trace_c2.npz"])
def test_fp64_full_frame_matches_oracle(acc, golden, fixture):
    d = golden(fixture)
    cam, m = _setup(d)
    o = O.default_opts(abi.PRECISION_FP64)
    o.n_bins = int(d["n_bins"])
    print(f"[case] {fixture} fp64 {cam.width}x{cam.height}", flush=True)
    out, st = acc.trace_batch([cam], [m], o)
    got = out.take(d["pix"])
    band = np.abs(d["crit"]) > 0.5
    ang, rel, rd = _compare(got, d, band, fp64=True)
    assert ang < 1e-6 and rel < 1e-6 and rd < 1e-6, (ang, rel, rd)
    assert np.all(out.cls <= abi.RAY_DISK)
    assert st["steps_bulk"] == out.steps.astype(np.int64).sum()
    assert out.null_drift.max() < 2e-6
    # the statistics pass's drift maximum and median (1/16-decade histogram) against the outputs
    assert st["null_drift_max"] == float(out.null_drift.max())
    med = float(np.median(out.null_drift.astype(np.float64)))
    if med > 1e-19:
        assert abs(np.log10(st["null_drift_median"]) - np.log10(med)) <= 1.0 / 16 + 1e-3, (st["null_drift_median"], med)


@pytest.mark.parametrize("fixture", ["ref_trace_c1.npz", "ref_trace_c2.npz"])
def test_fp32_escalation_meets_the_bar(acc, golden, fixture):
    d = golden(fixture)
    cam, m = _setup(d)
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.n_bins = int(d["n_bins"])
    print(f"[case] {fixture} fp32 {cam.width}x{cam.height}", flush=True)
    out, st = acc.trace_batch([cam], [m], o)
    got = out.take(d["pix"])
    band = np.abs(d["crit"]) > 0.5
    ang, rel, rd = _compare(got, d, band, fp64=False)
    assert ang < 1e-4 and rel < 1e-3, (ang, rel)
    # rays gated into the FP64 band are those near the critical curve
    e = (got.flags & abi.FLAG_ESCALATED) != 0
    assert st["rays_escalated"] == int(((out.flags & abi.FLAG_ESCALATED) != 0).sum())
    assert st["rays_escalated"] < 0.15 * out.n


def test_full_frame_fp32_meets_the_bar_at_every_pixel(acc):
    """The 1920x1080 Kerr a = 0.9 frame with the library's default options (the measured
    configuration): the FP32 path against the plain FP64 tracer at all 2,073,600 pixels. The FP64
    frame's own agreement with the oracle is pinned on the golden pixels of the same frame
    (ref_trace_c2) above."""
    from sirius_amd.accelerator import default_opts
    cam = abi.camera(50.0, 75.0, 20.0, 1920, 1080, disk=(0.0, 30.0))
    m = abi.metric_ks(1.0, 0.9)
    print("[case] 1920x1080 fp64 and fp32 frames, every pixel", flush=True)
    ref, _ = acc.trace_batch([cam], [m], default_opts(abi.PRECISION_FP64))
    got, _ = acc.trace_batch([cam], [m], default_opts(abi.PRECISION_FP32_ESCALATE))
    crit = acc.critical_distance([cam], [m])
    keep = np.abs(crit) > 0.5
    assert keep.mean() > 0.9
    assert np.array_equal(got.cls[keep], ref.cls[keep])
    esc = keep & (got.cls == abi.RAY_ESCAPED)
    ang = _angle(got.direction()[esc], ref.direction()[esc])
    dk = keep & (got.cls == abi.RAY_DISK)
    Ir, If = ref.intensity[:, dk], got.intensity[:, dk]
    scale = np.median(Ir.max(0))
    rel = np.abs(If - Ir).max(0) / np.maximum(Ir.max(0), 1e-2 * scale)
    print(f"escaped {esc.sum()}, disk {dk.sum()}, max angle {ang.max():.3e} rad, "
          f"max intensity error {rel.max():.3e}", flush=True)
    assert esc.sum() > 20000 and dk.sum() > 1000000  # the view is mostly disk
    assert ang.max() < 1e-4 and rel.max() < 1e-3, (ang.max(), rel.max())


def test_schwarzschild_frame_meets_the_bar_at_every_pixel(acc, golden):
    """The 256x256 Schwarzschild frame (r_obs 30 M, 80 deg, FOV 30 deg, disk 6-20 M, 8 bins) with
    the library's default options: the FP32 path against the plain FP64 tracer at every pixel
    outside the 0.5 M critical band, to the contract's bars. The FP64 frame's own agreement
    with the oracle is pinned on the golden pixels of the same frame (ref_trace_c1) above."""
    from sirius_amd.accelerator import default_opts
    d = golden("ref_trace_c1.npz")
    cam, m = _setup(d)
    o64, o32 = default_opts(abi.PRECISION_FP64), default_opts(abi.PRECISION_FP32_ESCALATE)
    o64.n_bins = o32.n_bins = int(d["n_bins"])
    print("[case] 256x256 Schwarzschild fp64 and fp32 frames, every pixel", flush=True)
    ref, _ = acc.trace_batch([cam], [m], o64)
    got, _ = acc.trace_batch([cam], [m], o32)
    keep = np.abs(acc.critical_distance([cam], [m])) > 0.5
    assert np.array_equal(got.cls[keep], ref.cls[keep])
    esc = keep & (got.cls == abi.RAY_ESCAPED)
    ang = _angle(got.direction()[esc], ref.direction()[esc])
    dk = keep & (got.cls == abi.RAY_DISK)
    Ir, If = ref.intensity[:, dk], got.intensity[:, dk]
    rel = np.abs(If - Ir).max(0) / np.maximum(Ir.max(0), 1e-2 * np.median(Ir.max(0)))
    print(f"escaped {esc.sum()}, disk {dk.sum()}, max angle {ang.max():.3e} rad, "
          f"max intensity error {rel.max():.3e}", flush=True)
    assert esc.sum() > 1000 and dk.sum() > 1000
    assert ang.max() < 1e-4 and rel.max() < 1e-3, (ang.max(), rel.max())


def test_persistent_queue_equals_grid_launch(acc):
    """Same per-ray FP32 arithmetic under both schedules: the work queue traces every ray
    exactly once, so all outputs are bit-identical."""
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, 320, 180, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.escalate_margin = o.escalate_band = o.escalate_disk_edge = 0.0
    a, sa = acc.trace_batch([cam], [m], o)
    o.schedule = abi.SCHEDULE_GRID
    b, sb = acc.trace_batch([cam], [m], o)
    for k in ("cls", "theta", "phi", "lam", "disk_r", "redshift", "steps", "intensity"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    # the critical flag comes from the persistent schedule's gate (the grid schedule has none)
    assert np.array_equal(a.flags & ~np.uint8(abi.FLAG_CRITICAL), b.flags)
    assert ((a.flags & abi.FLAG_CRITICAL) != 0).any()
    assert sa["steps_bulk"] == sb["steps_bulk"]
    # refill thresholds and launch shapes do not change results either
    o.schedule = abi.SCHEDULE_PERSISTENT
    for bs, bpc, rf in ((64, 1, 1), (128, 2, 32), (256, 0, 64)):
        o.block_size, o.blocks_per_cu, o.refill_min = bs, bpc, rf
        c, _ = acc.trace_batch([cam], [m], o)
        assert np.array_equal(c.cls, a.cls) and np.array_equal(c.intensity, a.intensity)


def test_slot_reservation_traces_every_ray_once(acc):
    """Waves reserve work-list slots in chunks of 128 while the queue head is far from the list's
    end, and exactly their idle lanes near it. On a small grid (one 64-lane wave per CU at the
    first shape) most of a 640x360 frame is served from chunks. Every ray is traced exactly once
    under every shape: outputs and the total attempt count equal the grid schedule's."""
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, 640, 360, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.escalate_margin = o.escalate_band = o.escalate_disk_edge = 0.0
    o.schedule = abi.SCHEDULE_GRID
    print("[case] slot reservation 640x360, grid schedule", flush=True)
    ref, sref = acc.trace_batch([cam], [m], o)
    o.schedule = abi.SCHEDULE_PERSISTENT
    for bs, bpc, rf in ((64, 1, 1), (64, 1, 8), (256, 1, 16), (128, 2, 32), (64, 4, 64), (256, 0, 16)):
        o.block_size, o.blocks_per_cu, o.refill_min = bs, bpc, rf
        print(f"[case] slot reservation 640x360, block {bs} x {bpc} per CU, refill {rf}", flush=True)
        got, st = acc.trace_batch([cam], [m], o)
        assert st["steps_bulk"] == sref["steps_bulk"], (bs, bpc, rf)
        for k in ("cls", "theta", "phi", "lam", "disk_r", "steps", "intensity"):
            assert np.array_equal(getattr(got, k), getattr(ref, k)), (k, bs, bpc, rf)


def test_escalated_rays_are_bitwise_fp64(acc):
    """Escalated rays (the gate's critical band, in-flight turning points inside the photon
    region, disk crossings at the inner edge) are traced by the FP64 persistent kernel with the
    FP32 path's step caps and the escalation tolerance: they equal an all-FP64 run with that
    policy bit for bit."""
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, 480, 270, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.escalate_band, o.escalate_margin, o.escalate_disk_edge = 0.25, 1.0, 0.02  # every criterion on
    o.escalate_rtol = 1e-8
    got, st = acc.trace_batch([cam], [m], o)
    o64 = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o64.precision = abi.PRECISION_FP64
    o64.rtol, o64.atol = o.escalate_rtol, o.atol * (o.escalate_rtol / o.rtol)  # as the library forms it
    ref, _ = acc.trace_batch([cam], [m], o64)
    e = (got.flags & abi.FLAG_ESCALATED) != 0
    assert e.sum() > 10
    for k in ("cls", "theta", "phi", "lam", "disk_r", "redshift", "steps", "intensity"):
        assert np.array_equal(getattr(got, k)[..., e], getattr(ref, k)[..., e]), k
    crit = acc.critical_distance([cam], [m])
    gated = np.abs(crit) < o.escalate_band - 1e-3  # the gate decides in FP32
    assert np.all(e[gated])


def test_multi_camera_batch_equals_single_cameras(acc):
    """Cameras with different spacetimes in one persistent launch give exactly the results of
    tracing each camera alone."""
    metrics = [abi.metric_ks(1.0, 0.3, 0.2, 5e-5), abi.metric_ks(1.0, 0.95), abi.metric_ks(1.0, 0.0, 0.4)]
    cams = [abi.camera(25.0, 30.0, 25.0, 64, 48, metric_index=0),
            abi.camera(60.0, 85.0, 15.0, 37, 29, metric_index=1),
            abi.camera(40.0, 60.0, 35.0, 50, 50, metric_index=2)]
    crit = acc.critical_distance(cams, metrics)
    for prec in (abi.PRECISION_FP64, abi.PRECISION_FP32_ESCALATE):
        o = O.default_opts(prec)
        o.n_bins = 8
        batch, _ = acc.trace_batch(cams, metrics, o)
        off = 0
        for c in cams:
            single_cam = abi.Camera.from_buffer_copy(c)
            single_cam.metric_index = 0
            one, _ = acc.trace_batch([single_cam], [metrics[c.metric_index]], o)
            n = c.width * c.height
            sl = slice(off, off + n)
            if prec == abi.PRECISION_FP64:
                # FP64: the multi-camera and single-camera kernel variants agree bit for bit
                for k in ("cls", "theta", "phi", "disk_r", "steps"):
                    assert np.array_equal(getattr(batch, k)[sl], getattr(one, k)), k
                assert np.array_equal(batch.intensity[:, sl], one.intensity)
            else:
                # FP32: the two variants contract FMAs differently, so each is held to the FP32
                # bar against the camera's FP64 result
                band = np.abs(crit[sl]) > 0.5
                o64 = O.default_opts(abi.PRECISION_FP64)
                o64.n_bins = 8
                ref, _ = acc.trace_batch([single_cam], [metrics[c.metric_index]], o64)
                for got in (batch.take(np.arange(off, off + n)), one):
                    assert np.array_equal(got.cls[band], ref.cls[band])
                    esc = band & (got.cls == 0) & (ref.cls == 0)
                    assert _angle(got.direction()[esc], ref.direction()[esc]).max() < 1e-4
                    dk = band & (got.cls == 2) & (ref.cls == 2)
                    I1, I0 = got.intensity[:, dk], ref.intensity[:, dk]
                    rel = np.abs(I1 - I0).max(0) / np.maximum(I0.max(0), 1e-2 * np.median(I0.max(0)))
                    assert rel.max() < 1e-3
            off += n


def test_charged_and_lambda_spacetimes_vs_oracle(acc):
    """Kerr-Newman-de Sitter parameters (Q, Lambda) through the Sherman-Morrison inverse
    against the oracle's general inverse."""
    for M, a, Q, L in ((1.0, 0.6, 0.5, 0.0), (1.0, 0.2, 0.0, 1e-4), (1.0, 0.8, 0.3, 8e-5)):
        m = abi.metric_ks(M, a, Q, L)
        cam = abi.camera(40.0, 70.0, 30.0, 40, 30, disk=(0.0, 20.0))
        o = O.default_opts(abi.PRECISION_FP64)
        o.n_bins = 8
        ref, _ = O.trace([cam], [m], o)
        got, _ = acc.trace_batch([cam], [m], o)
        assert np.array_equal(got.cls, ref.cls)
        esc = ref.cls == 0
        assert _angle(got.direction()[esc], ref.direction()[esc]).max() < 1e-6
        dk = ref.cls == 2
        assert np.allclose(got.intensity[:, dk], ref.intensity[:, dk], rtol=1e-6, atol=1e-12)


def test_critical_distance_matches_oracle(acc):
    for a in (0.0, 0.9):
        m = abi.metric_ks(1.0, a)
        cam = abi.camera(50.0, 75.0, 20.0, 64, 36)
        got = acc.critical_distance([cam], [m])
        ref = np.array([O.critical_distance(cam, m, i % 64, i // 64) for i in range(64 * 36)])
        assert np.abs(got - ref).max() < 1e-4


def test_edge_cases_and_errors(acc):
    m = abi.metric_ks(1.0, 0.5)
    o = O.default_opts(abi.PRECISION_FP64)
    out, _ = acc.trace_batch([abi.camera(30.0, 60.0, 30.0, 1, 1)], [m], o)
    assert out.cls[0] == abi.RAY_CAPTURED
    o.max_steps = 5
    out, _ = acc.trace_batch([abi.camera(30.0, 60.0, 30.0, 3, 2)], [m], o)
    assert np.all(out.cls == abi.RAY_MAX_STEPS) and np.all(out.steps == 5)
    from sirius_amd.accelerator import SiriusHipError
    bad = O.default_opts(abi.PRECISION_FP64)
    bad.n_bins = 65
    with pytest.raises(SiriusHipError) as e:
        acc.trace_batch([abi.camera(30.0, 60.0, 30.0, 4, 4)], [m], bad)
    assert e.value.rc == abi.E_INVALID
    with pytest.raises(SiriusHipError):
        acc.trace_batch([abi.camera(30.0, 60.0, 30.0, 4, 4, metric_index=2)], [m], o)
    with pytest.raises(SiriusHipError):
        acc.trace_batch([abi.camera(30.0, 60.0, 30.0, 4, 4)], [abi.metric_ks(1.0, 1.2)], o)
    with pytest.raises(SiriusHipError):
        acc.trace_batch([abi.camera(30.0, 60.0, 30.0, 0, 4)], [m], o)


def test_accelerator_launch_shades_framebuffer(acc):
    """IAccelerator path: initialise -> launch -> getFrameBuffer gives a W x H RGBA image made
    from the physical outputs (disk pixels lit, shadow black)."""
    w, h = 64, 36
    acc.initialise(w, h)
    assert acc.isInitialised() and acc.getFrameBufferSize() == w * h * 16
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, w, h, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    acc.launch(cam, m, o)
    fb = acc.getFrameBuffer().reshape(h, w, 4)
    ref, _ = acc.trace_batch([cam], [m], o)
    cls = ref.cls.reshape(h, w)
    assert np.all(fb[..., 3] == 1.0)
    assert np.all(fb[cls == abi.RAY_CAPTURED][:, :3] == 0.0)
    assert fb[cls == abi.RAY_DISK][:, :3].max() > 0.0
    bg = np.zeros((8, 16, 4), np.uint8)
    bg[..., 1] = 200
    acc.uploadBackground(bg)
    acc.launch(cam, m, o)
    fb2 = acc.getFrameBuffer().reshape(h, w, 4)
    if (cls == abi.RAY_ESCAPED).any():
        assert np.allclose(fb2[cls == abi.RAY_ESCAPED][:, 1], 200 / 255)
    with pytest.raises(Exception):
        acc.launch(abi.camera(50.0, 75.0, 20.0, w + 1, h), m, o)


def _reference_background(theta, phi, tex):
    """RenderSession::sampleStarfield (SNRS001A.cpp:631-703) at the escape angles: phi into
    [0, 2 pi), u = phi / 2 pi, v = theta / pi, bilinear between the texels around
    (u (W - 1), v (H - 1)); coordinates in FP64, weights and sums in float."""
    h, w = tex.shape[:2]
    ph = phi.astype(np.float64)
    ph = np.where(ph < 0, ph + 2 * np.pi, ph)
    u, v = ph / (2 * np.pi), theta.astype(np.float64) / np.pi
    px, py = u * (w - 1), v * (h - 1)
    x0, y0 = np.floor(px).astype(int), np.floor(py).astype(int)
    x1, y1 = np.minimum(x0 + 1, w - 1), np.minimum(y0 + 1, h - 1)
    fx, fy = px - x0, py - y0
    w00 = ((1 - fx) * (1 - fy)).astype(np.float32)
    w10 = (fx * (1 - fy)).astype(np.float32)
    w01 = ((1 - fx) * fy).astype(np.float32)
    w11 = (fx * fy).astype(np.float32)
    c = tex[..., :3].astype(np.float32) / np.float32(255)
    out = c[y0, x0] * w00[:, None] + c[y0, x1] * w10[:, None]
    out = out + c[y1, x0] * w01[:, None]
    return out + c[y1, x1] * w11[:, None]


def test_background_is_the_reference_sampler(acc):
    """Escaped pixels of launch take the equirectangular background as the reference's session
    samples it, bilinear, to float rounding (the same float steps in the same order)."""
    w, h = 80, 45
    acc.initialise(w, h)
    tex = np.random.default_rng(9).integers(0, 256, (64, 128, 4), dtype=np.uint8)
    acc.uploadBackground(tex)
    m = abi.metric_ks(1.0, 0.6)
    cam = abi.camera(30.0, 70.0, 60.0, w, h, disk=(0.0, 0.5))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    o.emission, o.n_bins = abi.EMISSION_NONE, 0
    print("[case] background sampler", flush=True)
    acc.launch(cam, m, o)
    fb = acc.getFrameBuffer().reshape(-1, 4)[:, :3]
    rays, _ = acc.trace_batch([cam], [m], o)
    esc = rays.cls == abi.RAY_ESCAPED
    assert esc.sum() > 0.5 * w * h
    want = _reference_background(rays.theta[esc], rays.phi[esc], tex)
    assert np.abs(fb[esc] - want).max() <= 2e-7, np.abs(fb[esc] - want).max()
    assert np.all(fb[~esc] == 0)
    acc.uploadBackground(np.zeros((1, 1, 4), np.uint8))


def test_last_results_follow_the_library_buffers(acc):
    """fetch_results returns the last trace's results while the library's result arrays hold
    them: after launch they are the frame's rays; once a later call has replaced the arrays
    (here an autotune over a larger batch) fetching is an error, never a read of freed memory."""
    m = abi.metric_ks(1.0, 0.9)
    small = abi.camera(50.0, 75.0, 20.0, 32, 18, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    print("[case] last results and buffer replacement", flush=True)
    acc.trace_resident([small], [m], o)
    got = acc.fetch_results()
    ref, _ = acc.trace_batch([small], [m], o)
    assert np.array_equal(got.cls, ref.cls) and np.array_equal(got.theta, ref.theta)
    w, h = 40, 24
    acc.initialise(w, h)
    frame = abi.camera(50.0, 75.0, 20.0, w, h, disk=(0.0, 30.0))
    acc.launch(frame, m, o)
    after_launch = acc.fetch_results()
    want, _ = acc.trace_batch([frame], [m], o)
    assert after_launch.n == w * h and np.array_equal(after_launch.cls, want.cls)
    acc.trace_resident([small], [m], o)
    big = abi.camera(50.0, 75.0, 20.0, 640, 360, disk=(0.0, 30.0))
    acc.autotune([big], [m], O.default_opts(abi.PRECISION_FP32_ESCALATE))
    with pytest.raises(SiriusHipError) as e:
        acc.fetch_results()
    assert e.value.rc == abi.E_STATE


@pytest.mark.gpu
def test_results_in_caller_device_arrays_are_not_held(acc):
    """A trace into the caller's device arrays leaves only the library-owned fields as the held
    results: after the caller frees its arrays, fetch_results copies the library's own fields
    (here the state, not given by the caller) and leaves the caller's fields untouched, never
    reading the freed memory."""
    from devmem import DeviceArray
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, 48, 27, disk=(0.0, 30.0))
    n = 48 * 27
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    print("[case] caller-owned device outputs", flush=True)
    cls_d = DeviceArray(shape=(n,), dtype=np.uint8)
    theta_d = DeviceArray(shape=(n,), dtype=np.float32)
    ro = abi.RayOutput(on_device=1)
    ro.cls, ro.theta = cls_d.data_ptr(), theta_d.data_ptr()
    acc.trace_batch([cam], [m], o, device_out=ro)
    want_cls = cls_d.get()
    del cls_d, theta_d  # the caller frees its arrays
    got = acc.fetch_results(with_state=True)
    assert (got.cls == 0).all() and (got.theta == 0).all()  # not copied: the caller's fields
    ref, _ = acc.trace_batch([cam], [m], o, with_state=True)
    assert np.array_equal(want_cls, ref.cls)
    assert np.array_equal(got.state, ref.state) and np.array_equal(got.phi, ref.phi)


def _audit_selection(n, fraction, seed):
    """The precision audit's ray choice restated: splitmix64 of (seed ^ ray index), taken when it
    is at most fraction x 2^64 (every ray at fraction >= 1)."""
    ids = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) ^ ids) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    thr = (1 << 64) - 1 if fraction >= 1 else int(fraction * 18446744073709551615.0)
    return z <= np.uint64(thr)


def _chord_angle(ta, pa, tb, pb):
    a, b = _dirs(ta, pa), _dirs(tb, pb)
    return 2.0 * np.arcsin(np.minimum(1.0, 0.5 * np.linalg.norm(a - b, axis=1)))


@pytest.mark.parametrize("fraction,seed", [(0.05, 0x51121010), (1.0, 0)])
def test_precision_audit_matches_a_direct_fp64_comparison(acc, fraction, seed):
    """The per-call error estimate of the FP32 path (SiriusHipTraceOpts::precision_audit): the
    library re-traces a seeded fraction of the frame's rays with the plain FP64 integrator and
    reports class mismatches, the largest escape-direction and per-bin intensity errors and their
    medians, over rays outside the critical band (those the accuracy contract covers). Checked
    against the same comparison made here between the FP32 frame and the plain FP64 frame: the
    same rays chosen (the selection hash restated), the same counts, the maxima
    equal to float rounding, the medians within one quarter-decade histogram bin."""
    from sirius_amd.accelerator import default_opts
    w, h = 480, 270
    cam = abi.camera(50.0, 75.0, 20.0, w, h, disk=(0.0, 30.0))
    m = abi.metric_ks(1.0, 0.9)
    print(f"[case] precision audit fraction {fraction} seed {seed:#x} {w}x{h}", flush=True)
    ref, _ = acc.trace_batch([cam], [m], default_opts(abi.PRECISION_FP64))
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    o.precision_audit, o.audit_seed = fraction, seed
    got, st = acc.trace_batch([cam], [m], o)
    crit = (got.flags & abi.FLAG_CRITICAL) != 0
    assert crit.sum() > 1000  # the band is skipped: the audit estimates the contract's error
    sel = _audit_selection(w * h, fraction, seed) & ~crit
    assert sel.sum() > 1000
    assert st["precision_audited"] == int(sel.sum())
    same = sel & (got.cls == ref.cls)
    assert st["precision_class_mismatch"] == int((sel & ~same).sum())
    esc = same & (ref.cls == abi.RAY_ESCAPED)
    de = _chord_angle(ref.theta[esc], ref.phi[esc], got.theta[esc], got.phi[esc])
    dk = same & (ref.cls == abi.RAY_DISK)
    Ir, If = ref.intensity[:, dk].astype(np.float64), got.intensity[:, dk].astype(np.float64)
    imax = Ir.max(0)
    ie = np.abs(If - Ir).max(0)[imax > 0] / imax[imax > 0]
    print(f"audited {st['precision_audited']}, mismatches {st['precision_class_mismatch']}, "
          f"dir max {st['precision_dir_err_max']:.3e} (direct {de.max():.3e}), "
          f"int max {st['precision_int_err_max']:.3e} (direct {ie.max():.3e}), "
          f"medians {st['precision_dir_err_median']:.2e} {st['precision_int_err_median']:.2e}", flush=True)
    assert esc.sum() > 30 and ie.size > 1000  # the view is mostly disk
    assert st["precision_dir_err_max"] == pytest.approx(de.max(), rel=1e-5, abs=1e-15)
    assert st["precision_int_err_max"] == pytest.approx(ie.max(), rel=1e-5, abs=1e-15)

    def median_bin(e):
        b = np.clip(((np.log10(np.maximum(e, 1e-300)) + 12.0) * 4.0).astype(int), 0, 63)
        b[e <= 0] = 0
        return np.sort(b)[(len(b) - 1) // 2]

    for key, e in (("precision_dir_err_median", de), ("precision_int_err_median", ie)):
        got_bin = (np.log10(st[key]) + 12.0) * 4.0 - 0.5
        assert abs(got_bin - median_bin(e)) <= 1.0 + 1e-9, (key, st[key], got_bin, median_bin(e))
    assert st["precision_audit_ms"] >= 0.0
    # the FP32 results themselves are those of a frame without the audit
    plain, _ = acc.trace_batch([cam], [m], default_opts(abi.PRECISION_FP32_ESCALATE))
    assert np.array_equal(plain.cls, got.cls) and np.array_equal(plain.theta, got.theta)
    assert np.array_equal(plain.flags, got.flags)
