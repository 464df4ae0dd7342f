"""Benchmark: Mrays/s at 1920x1080, Kerr a = 0.9 (observer r = 50 M, inclination 75 deg,
FOV 20 deg, Novikov-Thorne disk from the ISCO to 30 M, 16 spectral bins).

One step = one full frame traced on the GPU with every result left in device memory.
  value          the FP32 path (compensated FP32 state, FP64 null-constraint re-normalisation,
                 escalation criteria at their defaults)
  fp64_plain     the plain FP64 integrator on the same frame (the comparison baseline)
  fp64_same_step_policy
                 the FP64 integrator at the FP32 path's step policy (rtol, far-field cap,
                 inward step proposal): the part of the speed-up that precision alone buys
  precision_audit
                 the per-call FP64 error estimate: 1 % of the frame's rays re-traced by the plain
                 FP64 integrator inside the call, with class mismatches and error maxima/medians
  symplectic     the frame by the fixed-step symplectic integrator (h = 0.01 M, FP64) with its
                 null-constraint drift
  accuracy       the FP32 frame against the FP64 frame, pixel by pixel, outside the critical band
                 and (reported apart) inside it; per-pixel and per-bin intensity errors
  with_edge_escalation
                 the same frame with inner-edge disk crossings re-traced in FP64
  with_band_escalation
                 the same frame with the rays within 0.5 M of the critical curve traced in FP64
                 on a second stream, concurrently with the FP32 tracer
  roofline       dominant kernel (the FP32 persistent tracer): algorithmic flops per launch
                 (1,326 per Kerr-Schild DP45 attempt) over its HIP-event time, against the
                 FP32 vector peak; traffic from the profile's HBM byte counters
  cpu_baseline   the reference's own CPU integrator (oracle/_ref, compiled from its sources)
                 on every pixel of the frame, on this host's cores (about 15 s), and the
                 reference's image writers on the shaded frame
  other_workloads
                 the 64-camera random batch (volumetric, autotuned launch) and the wormhole and
                 warp scenes at 1024x1024: Mrays/s, geodesic steps/s, null-constraint drift
  surrogate      the MLP surrogate as C4 configures it (one-time training on 2^24 integrator-
                 labelled rays, excluded from the frame time and reported apart): the error gate's
                 coverage against its threshold, the headline frame with it on, a sky-dominated
                 view on and off, and the fused inference kernel's throughput

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-accuracy]
                [--no-surrogate] [--no-symplectic] [--no-extra] [--copy-calibration MIB]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s 1080p Kerr a=0.9: plain FP64 vs FP32+surrogate at ≤1e-4 rad error"
FP32_VALU_PEAK_TFLOPS = 157.3
FP64_VALU_PEAK_TFLOPS = 78.6
# HBM bytes per launch of the FP32 bulk tracer on this frame, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by a copy kernel of known size (profiles/trace_kerr_a0.9_1080p_counters.md)
BULK_HBM_BYTES = 221.0e6
W, H = 1920, 1080


def workload():
    from sirius_amd import abi
    metric = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, W, H, disk=(0.0, 30.0))
    return cam, metric


def calibration_copy(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (its copy kernel reads and
    writes exactly nbytes), for scaling the HBM byte counters of a profile."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), nbytes) == 0 and hip.hipMalloc(C.byref(b), nbytes) == 0
    assert hip.hipMemcpy(b, a, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def time_frames(acc, cam, metric, opts, steps, warmup, n_stats=3, multi=False):
    """Wall time per frame over `steps` frames (no statistics pass inside the timed region),
    then `n_stats` frames with statistics for the HIP-event kernel times and step counts.
    multi: cam and metric are lists (one batch of several cameras)."""
    cams, metrics = (cam, metric) if multi else ([cam], [metric])
    for _ in range(warmup):
        acc.trace_resident(cams, metrics, opts, stats=False)
    acc.synchronise()
    t0 = time.perf_counter()
    for _ in range(steps):
        acc.trace_resident(cams, metrics, opts, stats=False)
    acc.synchronise()
    dt = time.perf_counter() - t0
    stats = [acc.trace_resident(cams, metrics, opts) for _ in range(n_stats)]
    return dt / steps, stats


def tune_fp64(acc, cam, metric, o):
    """Block size of the plain FP64 tracer (its grid schedule), the fastest of {64, 128, 256} by
    the faster of two timed frames each, so that both sides of the FP32/FP64 ratio run a tuned
    launch. Runs before the timed region."""
    best = None
    for bs in (64, 128, 256):
        o.block_size = bs
        ms = min(acc.trace_resident([cam], [metric], o)["kernel_ms"] for _ in range(2))
        if best is None or ms < best[0]:
            best = (ms, bs)
    o.block_size = best[1]
    return o


def _angles(u, v):
    return np.arccos(np.clip((u * v).sum(1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1), -1, 1))


def _intensity_errors(If, Ir, scale, prefix, out):
    """Intensity error of disk pixels (bins x pixels): per pixel, the largest bin error over the
    pixel's brightest bin (floored at 1e-2 of the frame's median brightest bin); and per bin, the
    error relative to that bin's own value, over bins above 1e-6 of the same median."""
    rel = np.abs(If - Ir).max(0) / np.maximum(Ir.max(0), 1e-2 * scale)
    out[prefix + "intensity_rel_err_max"] = float(rel.max())
    out[prefix + "intensity_rel_err_p999"] = float(np.percentile(rel, 99.9))
    out[prefix + "intensity_l1_rel_err"] = float(np.abs(If - Ir).sum() / Ir.sum())
    live = Ir > 1e-6 * scale
    if live.any():
        pb = np.abs(If - Ir)[live] / Ir[live]
        out[prefix + "intensity_per_bin_rel_err_max"] = float(pb.max())
        out[prefix + "intensity_per_bin_rel_err_p999"] = float(np.percentile(pb, 99.9))
        out[prefix + "intensity_per_bin_rel_err_median"] = float(np.median(pb))


def accuracy(fast, ref, crit, band=0.5):
    """The FP32 frame against the FP64 frame, pixel by pixel: outside the critical band (the
    contract's classes) and, reported apart, inside it (rays within `band` M of the critical
    curve: their direction and intensity errors where both frames agree on the class)."""
    out = {}
    keep = np.abs(crit) > band
    out["class_agreement_outside_band"] = float((fast.cls == ref.cls)[keep].mean())
    out["band_fraction"] = float((~keep).mean())
    out["class_agreement_in_band"] = float((fast.cls == ref.cls)[~keep].mean()) if (~keep).any() else 1.0
    dk_all = (fast.cls == 2) & (ref.cls == 2)
    scale = float(np.median(ref.intensity[:, dk_all].max(0))) if dk_all.any() else 1.0
    for prefix, region in (("", keep), ("band_", ~keep)):
        esc = region & (fast.cls == 0) & (ref.cls == 0)
        if esc.any():
            ang = _angles(fast.direction()[esc], ref.direction()[esc])
            out[prefix + "escape_dir_err_max_rad"] = float(ang.max())
            out[prefix + "escape_dir_err_p99_rad"] = float(np.percentile(ang, 99))
            out[prefix + "escaped_rays"] = int(esc.sum())
        dk = region & dk_all
        if dk.any():
            _intensity_errors(fast.intensity[:, dk], ref.intensity[:, dk], scale, prefix, out)
            out[prefix + "disk_rays"] = int(dk.sum())
    return out


def other_workloads(acc, steps, warmup):
    """The batched and non-Kerr workloads, one GPU each, FP32 path: the 64-camera random
    Kerr-Newman-de Sitter batch (volumetric emission, 8 bins, launch shape from the autotuner) and
    the Morris-Thorne and Alcubierre scenes at 1024x1024, with geodesic steps/s and the largest
    null-constraint drift |g^{mu nu} p_mu p_nu| of the final states."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from workloads import random_spacetime_batch, warp_scene, wormhole_scene
    from sirius_amd import abi
    from sirius_amd.accelerator import default_opts

    def measure(cams, metrics, o, n_steps, n_warm):
        n = sum(c.width * c.height for c in cams)
        ms, st = time_frames(acc, cams[0] if len(cams) == 1 else cams, metrics[0] if len(metrics) == 1 else metrics,
                             o, n_steps, n_warm, multi=len(cams) > 1)
        s = st[-1]
        k = float(np.mean([x["kernel_ms"] for x in st]))
        return {"rays": n, "ms_per_step": ms * 1e3, "mrays_per_s": n / ms / 1e6,
                "geodesic_steps_per_s": (s["steps_bulk"] + s["steps_escalated"]) / (ms),
                "steps_per_ray": s["steps_bulk"] / n, "rays_escalated": s["rays_escalated"],
                "null_drift_max": s["null_drift_max"], "null_drift_median": s["null_drift_median"],
                "tracer_kernel_ms": k}

    out = {}
    cams, metrics = random_spacetime_batch()
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    o.emission = abi.EMISSION_VOLUMETRIC
    o.n_bins = 8
    acc.autotune(cams, metrics, o)
    r = measure(cams, metrics, o, max(1, steps // 2), 1)
    r["launch"] = {"block_size": o.block_size, "blocks_per_cu": o.blocks_per_cu, "refill_min": o.refill_min}
    r["workload"] = "64 random Kerr-Newman-de Sitter cameras, volumetric, 8 bins"
    out["batch_64_cameras"] = r
    for name, (cam, metric) in (("wormhole_1024", wormhole_scene(1024)), ("warp_1024", warp_scene(1024))):
        o = default_opts(abi.PRECISION_FP32_ESCALATE)
        o.emission = abi.EMISSION_NONE
        out[name] = measure([cam], [metric], o, steps, warmup)
    out["wormhole_1024_two_spheres"] = wormhole_spheres(acc, steps)
    out["image_writers_1080p"] = image_writers(acc, steps)
    out["ray_bundle_1080p"] = ray_bundle(acc, steps)
    return out


def ray_bundle(acc, steps):
    """The 1080p Kerr frame's ray bundles (geodesic deviation in FP64, results left in device
    memory), and launch with the default star catalogue (100000 candidates, magnitude limit 12)
    point-sampled and averaged over each pixel's beam."""
    from sirius_amd import abi
    from sirius_amd.accelerator import default_opts, starfield_default, starfield_generate
    cam, metric = workload()
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    acc.trace_bundle([cam], [metric], o, resident=True)  # warm
    ts = []
    for _ in range(max(2, steps // 2)):
        t0 = time.perf_counter()
        acc.trace_bundle([cam], [metric], o, resident=True)
        ts.append(time.perf_counter() - t0)
    ms = float(np.mean(ts)) * 1e3
    out = {"bundle_ms": ms, "bundle_mrays_per_s": W * H / ms / 1e3,
           "tolerances": "the FP32 path's (rtol 1e-6, far-field cap 0.25 r), stepped in FP64"}
    cfg = starfield_default()
    stars, n = starfield_generate(cfg)
    acc.initialise(W, H)
    acc.set_starfield(0, cfg, stars, n)
    try:
        for name, on in (("launch_point_ms", False), ("launch_beam_ms", True)):
            acc.set_ray_bundle(on)
            acc.launch(cam, metric, o)
            ts = []
            for _ in range(max(2, steps // 2)):
                t0 = time.perf_counter()
                acc.launch(cam, metric, o)
                ts.append(time.perf_counter() - t0)
            out[name] = float(np.mean(ts)) * 1e3
        out["stars"] = n
    finally:
        acc.set_ray_bundle(False)
        acc.set_starfield(0, None, None, 0)
    return out


def wormhole_spheres(acc, steps):
    """The wormhole scene at 1024x1024 rendered through launch with its two celestial spheres:
    4096 stars each, magnitudes ~ Exp(1), seeds 0x51121008 (near side) and 0x51121009 (far side)
    (tests/workloads.py): the trace, the star light of every escaped pixel from its side's
    sphere, and the framebuffer."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from workloads import WORMHOLE_SPHERE_SEEDS, sphere_catalogue, wormhole_scene
    from sirius_amd import abi
    from sirius_amd.accelerator import default_opts, starfield_default
    cam, metric = wormhole_scene(1024)
    o = default_opts(abi.PRECISION_FP32_ESCALATE)
    o.emission = abi.EMISSION_NONE
    cfg = starfield_default()
    acc.initialise(1024, 1024)
    try:
        for side, seed in enumerate(WORMHOLE_SPHERE_SEEDS):
            rows = sphere_catalogue(seed)
            stars = (abi.Star * len(rows))()
            np.frombuffer(stars, np.float32).reshape(-1, 8)[:] = rows
            acc.set_starfield(side, cfg, stars, len(rows))
        acc.launch(cam, metric, o)
        ts = []
        for _ in range(max(2, steps // 2)):
            t0 = time.perf_counter()
            acc.launch(cam, metric, o)
            ts.append(time.perf_counter() - t0)
        ms = float(np.mean(ts)) * 1e3
        return {"launch_ms": ms, "mrays_per_s": 1024 * 1024 / ms / 1e3, "stars_per_side": 4096}
    finally:
        acc.set_starfield(0, None, None, 0)
        acc.set_starfield(1, None, None, 0)


# Algorithmic HBM bytes per pixel of each format's device conversion: the float RGBA frame read
# (16 B) plus the payload written; PNG also writes and re-reads its 8-bit pixels for the row
# filters (4 + 4 B) before writing the filtered rows (4 B).
IMAGE_BYTES_PER_PIXEL = {"png": 16 + 4 + 4 + 4, "exr": 16 + 8, "ppm": 16 + 3}


def image_writers(acc, steps):
    """The 1080p Kerr frame written to files as PNG, EXR and PPM from the framebuffer in device
    memory: the device conversion (HIP events), the copy of the payload to pinned host memory, the
    host compression on up to 16 threads and the whole call, per format (the reference's own
    writers are timed on the same frame by the CPU-baseline leg)."""
    from sirius_amd import abi
    from sirius_amd.accelerator import default_opts
    cam, metric = workload()
    acc.initialise(W, H)
    acc.launch(cam, metric, default_opts(abi.PRECISION_FP32_ESCALATE))
    out = {"frame": "1920x1080 RGBA float (the Kerr a=0.9 benchmark frame, shaded)"}
    fmts = (("png", abi.IMAGE_PNG), ("exr", abi.IMAGE_EXR), ("ppm", abi.IMAGE_PPM))
    import tempfile
    tmp = tempfile.mkdtemp()
    for name, fmt in fmts:
        path = os.path.join(tmp, "frame." + name)
        acc.write_image(path, fmt)  # warm: staging buffers and threads
        # the whole call, file written, as the reference writers' time includes theirs
        sts = [acc.write_image(path, fmt) for _ in range(max(3, steps))]
        mean = lambda k: float(np.mean([getattr(x, k) for x in sts]))
        conv = mean("convert_ms")
        out[name] = {"total_ms": mean("total_ms"), "convert_ms": conv, "copy_ms": mean("copy_ms"),
                     "pack_ms": mean("pack_ms"), "file_bytes": int(sts[-1].file_bytes),
                     "convert_GBps": IMAGE_BYTES_PER_PIXEL[name] * W * H / (conv * 1e-3) / 1e9}
    return out


def reference_writers(acc, writers):
    """CPU baseline of the image writers: the reference's own PNGWriter::writeRGBA,
    EXRWriter::writeEXR and BufferWriter::writePPM (oracle/_ref, compiled in place) on the shaded
    1080p frame, one host core each, beside this backend's whole call."""
    from oracle import ref_py as RF
    from sirius_amd import abi
    from sirius_amd.accelerator import default_opts
    if not os.path.exists(RF.IMAGE_LIB_PATH):
        return {"unavailable": "oracle/_ref image library not built"}
    cam, metric = workload()
    acc.initialise(W, H)
    acc.launch(cam, metric, default_opts(abi.PRECISION_FP32_ESCALATE))
    frame = acc.getFrameBuffer().reshape(H, W, 4).copy()
    out = {}
    for name, fn in (("png", RF.write_png), ("exr", RF.write_exr), ("ppm", lambda f: RF.write_ppm_aces(f, 1.0))):
        t0 = time.perf_counter()
        fn(frame)
        ms = (time.perf_counter() - t0) * 1e3
        out[name] = {"reference_cpu_ms": ms}
        if name in writers:
            out[name]["speedup_of_backend"] = ms / writers[name]["total_ms"]
    return out


def _surrogate_calibration(acc):
    import ctypes as C
    L = acc._L
    L.sirius_hip_surrogate_calibration.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                                   C.POINTER(C.c_int)]
    err, n, bad = np.zeros(1024, np.float32), np.zeros(1024, np.uint32), np.zeros(1024, np.uint32)
    nb = C.c_int()
    L.sirius_hip_surrogate_calibration(acc._h, err.ctypes.data, n.ctypes.data, bad.ctypes.data, 1024, C.byref(nb))
    return err[:nb.value], n[:nb.value], bad[:nb.value]


def surrogate_ablation(acc, cam, metric, ref, steps, warmup):
    """The surrogate as C4 configures it: trained once on 2^24 integrator-labelled rays (4113
    cameras of 85 x 48 - the frame's 16:9 aspect - a ~ U[0, 0.99], inclination ~ U[5, 89] deg,
    r_obs ~ U[20, 100] M, FOV 20 deg; training and its held-out calibration excluded from the
    frame time, reported apart).
    Then: the error gate's coverage against its threshold on the held-out rays; the headline frame
    with the surrogate on; a sky-dominated view (r_obs 80 M, inclination 85 deg) on and off; and
    the fused inference kernel's throughput."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from workloads import MT19937_64
    from sirius_amd import abi
    from sirius_amd.accelerator import default_opts
    g = MT19937_64(0x51121006)
    cams, metrics = [], []
    n_train_cam, tw, th = 4113, 85, 48  # 2^24 rays and more, the frame's 16:9 aspect
    for i in range(n_train_cam):
        a, incl, r_obs = g.uniform(0.0, 0.99), g.uniform(5.0, 89.0), g.uniform(20.0, 100.0)
        metrics.append(abi.metric_ks(1.0, a))
        cams.append(abi.camera(r_obs, incl, 20.0, tw, th, disk=(0.0, 30.0), metric_index=i))
    lab = default_opts(abi.PRECISION_FP64)
    acc.mlp_init(0x51121006)
    epochs = 200
    t0 = time.perf_counter()
    loss = acc.surrogate_train(cams, metrics, lab, epochs, 8192, 2e-3, 0x51121006)
    train_s = time.perf_counter() - t0
    err, n, bad = _surrogate_calibration(acc)
    curve = []
    for tau in (2e-5, 5e-5, 1e-4, 3e-4, 1e-3):
        ok = (n >= 20) & (bad == 0) & (err <= tau)
        curve.append({"err_max_rad": tau, "bins": int(ok.sum()), "held_out_answers_kept": int(n[ok].sum())})
    out = {"train_rays": n_train_cam * tw * th, "train_cameras": f"{n_train_cam} x {tw}x{th}, FOV 20 deg",
           "train_epochs": epochs, "train_s": train_s, "train_final_loss": loss,
           "held_out_answers": int(n.sum()), "held_out_misclassified": int(bad.sum()),
           "coverage_vs_err_max": curve, "err_max_default": lab.surrogate_err_max}

    def frame(cam_, metric_, use, err_max=None):
        o = default_opts(abi.PRECISION_FP32_ESCALATE)
        o.use_surrogate = use
        o.audit_fraction = 0.0
        if err_max is not None:
            o.surrogate_err_max = err_max
        ms, st = time_frames(acc, cam_, metric_, o, steps, warmup)
        return ms, st[-1]

    def headline(err_max=None):
        ms, st = frame(cam, metric, 1, err_max)
        got = acc.fetch_results()
        sur = (got.flags & abi.FLAG_SURROGATE) != 0
        r = {"ms_per_step": ms * 1e3, "value": W * H / ms / 1e6, "answered": int(sur.sum()),
             "candidates": st["rays_candidates"], "gated": st["rays_gated"],
             "surrogate_ms": st["surrogate_ms"], "escaped_rays": int((ref.cls == 0).sum())}
        if sur.any():
            u, v = got.direction()[sur], ref.direction()[sur]
            u, v = u / np.linalg.norm(u, axis=1, keepdims=True), v / np.linalg.norm(v, axis=1, keepdims=True)
            e = 2 * np.arcsin(np.minimum(np.linalg.norm(u - v, axis=1) / 2, 1))
            r["class_mismatch_answered"] = int((ref.cls[sur] != got.cls[sur]).sum())
            r["escape_dir_err_max_rad"] = float(e.max())
            r["escape_dir_err_median_rad"] = float(np.median(e))
        return r

    out["headline"] = headline()
    # the gate at the bar itself (bins whose held-out error is at most 1e-4 rad answer)
    out["headline_err_max_1e-4"] = headline(1e-4)
    sky_cam, sky_metric = abi.camera(80.0, 85.0, 20.0, W, H, disk=(0.0, 30.0)), abi.metric_ks(1.0, 0.9)
    ms_off, _ = frame(sky_cam, sky_metric, 0)
    ms_on, st_on = frame(sky_cam, sky_metric, 1)
    out["sky_view"] = {"workload": "1920x1080 Kerr a=0.9, r_obs 80 M, inclination 85 deg, FOV 20 deg",
                       "off_ms": ms_off * 1e3, "on_ms": ms_on * 1e3, "answered": st_on["rays_surrogate"],
                       "candidates": st_on["rays_candidates"], "surrogate_ms": st_on["surrogate_ms"]}
    out["fused_inference"] = fused_inference(acc, 1 << 21)
    return out


def fused_inference(acc, rows, reps=10):
    """The fused inference kernel on `rows` random inputs with the current network: wall time per
    call over `reps` calls after one warm-up, and its BF16 MFMA rate (398,848 flops per row)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from devmem import DeviceArray
    xd = DeviceArray(np.random.default_rng(1).uniform(-1, 1, (rows, 5)).astype(np.float32))
    yd = DeviceArray(shape=(rows, 6), dtype=np.float32)
    acc.mlp_forward(xd.data_ptr(), yd.data_ptr(), rows)
    acc.synchronise()
    t0 = time.perf_counter()
    for _ in range(reps):
        acc.mlp_forward(xd.data_ptr(), yd.data_ptr(), rows)
    acc.synchronise()
    dt = (time.perf_counter() - t0) / reps
    return {"rows": rows, "ms": dt * 1e3, "tflops": rows * 398848 / dt / 1e12,
            "frac_of_bf16_peak": rows * 398848 / dt / 1e12 / 2500.0}


def symplectic_leg(acc, cam, metric):
    """The frame by the fixed-step symplectic integrator (Tao's extended phase space, Strang
    split, Yoshida-4 composition, h = 0.01 M, FP64), one frame after one warm-up: Mrays/s,
    geodesic steps/s and the null-constraint drift of the final states (max and median). Escaped
    rays run to r = 1000 M, about 1e5 fixed steps each."""
    from sirius_amd import abi
    from sirius_amd.accelerator import default_opts
    o = default_opts(abi.PRECISION_FP64)
    o.integrator = abi.INTEGRATOR_SYMPLECTIC
    o.symplectic_h = 0.01
    o.symplectic_order = 4
    o.max_steps = 400000
    acc.trace_resident([cam], [metric], o, stats=False)
    acc.synchronise()
    t0 = time.perf_counter()
    acc.trace_resident([cam], [metric], o, stats=False)
    acc.synchronise()
    dt = time.perf_counter() - t0
    st = acc.trace_resident([cam], [metric], o)
    got = acc.fetch_results()
    n = W * H
    return {"integrator": "Tao extended phase space (omega 20), Strang split, Yoshida-4, fixed h = 0.01 M, FP64",
            "ms_per_step": dt * 1e3, "mrays_per_s": n / dt / 1e6,
            "geodesic_steps_per_s": st["steps_bulk"] / dt, "steps_per_ray": st["steps_bulk"] / n,
            "kernel_ms": st["kernel_ms"], "achieved_tflops": st["flops"] / (st["kernel_ms"] * 1e-3) / 1e12,
            "flops_per_step": st["flops"] / max(st["steps_bulk"], 1),
            "null_drift_max": st["null_drift_max"], "null_drift_median": st["null_drift_median"],
            "class_counts": {k: int((got.cls == v).sum()) for k, v in
                             (("escaped", 0), ("captured", 1), ("disk", 2), ("max_steps", 3), ("failed", 5))}}


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on, within the
    OMP_NUM_THREADS share the machine allots (a GPU box gives one GPU 16 host cores)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(share))) if share.isdigit() and int(share) > 0 else n


def cpu_baseline(n_threads):
    """The reference's own integrator (oracle/_ref) when present, else the FP64 port, on every
    pixel of the frame."""
    from oracle import ref_py
    pix = np.arange(0, W * H, dtype=np.int64)
    if ref_py.available():
        t0 = time.perf_counter()
        ref_py.trace(1.0, 0.9, 0.0, 0.0, 50.0, np.deg2rad(75.0), 0.0, 20.0, W, H, pix,
                     r_cap=1.01 * (1 + np.sqrt(1 - 0.81)), r_esc=1000.0, disk=(ref_py.isco(0.9), 30.0),
                     threads=n_threads)
        dt = time.perf_counter() - t0
        kind = "reference"
    else:
        from oracle import oracle_py as O
        from sirius_amd import abi
        cam, metric = workload()
        o = O.default_opts(abi.PRECISION_FP64)
        t0 = time.perf_counter()
        O.trace([cam], [metric], o, ray_ids=pix, threads=n_threads)
        dt = time.perf_counter() - t0
        kind = "port"
    return {"value": len(pix) / dt / 1e6, "unit": "Mrays/s", "cores": n_threads, "kind": kind,
            "sample": f"every pixel of the 1920x1080 Kerr a=0.9 frame ({len(pix)} rays), "
                      f"FP64 DP45 rtol 1e-8, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-symplectic", action="store_true", help="skip the fixed-step symplectic frame")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and non-Kerr workloads")
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch shapes instead of tuned ones for the headline frame")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a copy kernel of known size that "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs scale against)")
    ap.add_argument("--mlp-only", type=int, default=0, metavar="ROWS",
                    help="time only the surrogate's fused inference kernel on ROWS random rows (a freshly "
                         "initialised network) and print that; for profiling the kernel alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this backend runs on one GPU; --gpus {args.gpus} is not supported "
              "(rays would shard across GPUs by camera or tile; see the design notes)", file=sys.stderr)
        sys.exit(2)

    from sirius_amd import abi
    from sirius_amd.accelerator import HipAccelerator, default_opts
    acc = HipAccelerator()
    if args.copy_calibration > 0:
        calibration_copy(args.copy_calibration << 20)
    if args.mlp_only > 0:
        acc.mlp_init(0x51121006)
        print(json.dumps({"fused_inference": fused_inference(acc, args.mlp_only)}))
        acc.close()
        return
    cam, metric = workload()
    n_rays = W * H

    o64 = default_opts(abi.PRECISION_FP64)
    o32 = default_opts(abi.PRECISION_FP32_ESCALATE)
    if not args.no_tune:
        # launch shapes picked before the timed region: the library's autotuner for the FP32 path
        # (block size x refill threshold at full occupancy, the whole frame timed), a block-size
        # sweep for the FP64 tracer
        tune_fp64(acc, cam, metric, o64)
        acc.autotune([cam], [metric], o32)
    ms64, st64 = time_frames(acc, cam, metric, o64, max(1, args.steps // 2), 1)
    ref = acc.fetch_results() if not args.no_accuracy else None

    ms32, st32 = time_frames(acc, cam, metric, o32, args.steps, args.warmup)
    fast = acc.fetch_results() if not args.no_accuracy else None

    s = st32[-1]
    kern_ms = float(np.mean([x["kernel_ms"] for x in st32]))
    achieved = s["flops"] / (kern_ms * 1e-3) / 1e12
    s64 = st64[-1]
    k64 = float(np.mean([x["kernel_ms"] for x in st64]))
    result = {
        "metric": METRIC,
        "value": n_rays / ms32 / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (compensated state; f64 null-constraint re-normalisation and escalation)",
        "data": "synthetic",
        "config": {"workload": "kerr_a0.9_1080p_dp45_fp32_escalate", "width": W, "height": H, "a": 0.9,
                   "r_obs": 50.0, "inclination_deg": 75.0, "fov_deg": 20.0, "disk": "ISCO..30M",
                   "n_bins": 16, "rtol": o32.rtol, "far_slope": o32.far_slope,
                   "escalate_band": o32.escalate_band, "escalate_margin": o32.escalate_margin,
                   "escalate_disk_edge": o32.escalate_disk_edge, "critical_band": o32.critical_band,
                   "surrogate": False},
        "roofline": {"bound": "valu", "kernel": "k_trace_persistent32<0,1,false>", "achieved": achieved,
                     "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_VALU_PEAK_TFLOPS,
                     "traffic": BULK_HBM_BYTES, "traffic_unit": "bytes per launch",
                     "traffic_source": "profiles/trace_kerr_a0.9_1080p_counters.md",
                     "flops_per_launch": s["flops"], "avg_launch_ms": kern_ms},
        "geodesic_steps_per_s": s["steps_bulk"] / (kern_ms * 1e-3),
        "rays_escalated": s["rays_escalated"],
        "escalate_ms": float(np.mean([x["escalate_ms"] for x in st32])),
        "fp64_plain": {"value": n_rays / ms64 / 1e6, "unit": "Mrays/s", "ms_per_step": ms64 * 1e3,
                       "kernel_ms": k64, "steps_per_ray": s64["steps_bulk"] / n_rays,
                       "achieved_tflops": s64["flops"] / (k64 * 1e-3) / 1e12,
                       "frac_of_fp64_peak": s64["flops"] / (k64 * 1e-3) / 1e12 / FP64_VALU_PEAK_TFLOPS},
        "speedup_vs_fp64_plain": ms64 / ms32,
        "null_drift_max": s["null_drift_max"],
        "null_drift_median": s["null_drift_median"],
        "launch": {"fp32": {"block_size": o32.block_size, "blocks_per_cu": o32.blocks_per_cu,
                            "refill_min": o32.refill_min},
                   "fp64": {"block_size": o64.block_size},
                   "tuned": not args.no_tune},
        "steps_per_ray": s["steps_bulk"] / n_rays,
    }
    # the plain FP64 tracer at the FP32 path's step policy (its rtol, atol, far-field cap and
    # inward step proposal): the share of the speed-up that comes from precision alone
    o64p = default_opts(abi.PRECISION_FP64)
    for k in ("rtol", "atol", "far_slope", "step_radial_exp", "block_size"):
        setattr(o64p, k, getattr(o32 if k != "block_size" else o64, k))
    ms64p, st64p = time_frames(acc, cam, metric, o64p, max(1, args.steps // 2), 1)
    result["fp64_same_step_policy"] = {
        "value": n_rays / ms64p / 1e6, "ms_per_step": ms64p * 1e3,
        "steps_per_ray": st64p[-1]["steps_bulk"] / n_rays, "speedup_of_fp32": ms64p / ms32,
        "policy": {"rtol": o64p.rtol, "atol": o64p.atol, "far_slope": o64p.far_slope,
                   "step_radial_exp": o64p.step_radial_exp}}
    # the per-call error estimate: 1 % of the frame's rays re-traced by the plain FP64 integrator
    # inside the same call and compared with the FP32 results (timed apart by HIP events)
    oa = default_opts(abi.PRECISION_FP32_ESCALATE)
    oa.block_size, oa.blocks_per_cu, oa.refill_min = o32.block_size, o32.blocks_per_cu, o32.refill_min
    oa.precision_audit, oa.audit_seed = 0.01, 0x51121010
    msa, sta = time_frames(acc, cam, metric, oa, max(1, args.steps // 2), 1)
    sa = sta[-1]
    result["precision_audit"] = {
        "fraction": oa.precision_audit, "ms_per_step": msa * 1e3,
        "audit_ms": float(np.mean([x["precision_audit_ms"] for x in sta])),
        "audited": sa["precision_audited"], "class_mismatch": sa["precision_class_mismatch"],
        "dir_err_max_rad": sa["precision_dir_err_max"], "dir_err_median_rad": sa["precision_dir_err_median"],
        "int_err_max": sa["precision_int_err_max"], "int_err_median": sa["precision_int_err_median"]}
    if not args.no_symplectic:
        result["symplectic"] = symplectic_leg(acc, cam, metric)
    if not args.no_accuracy:
        crit = acc.critical_distance([cam], [metric])
        result["accuracy_vs_fp64"] = accuracy(fast, ref, crit)
        # the same frame with the inner-edge escalation on: crossings within 2 % of r_in are
        # re-traced in FP64 after the bulk (the cost of that path, and what it buys)
        oe = default_opts(abi.PRECISION_FP32_ESCALATE)
        oe.block_size, oe.blocks_per_cu, oe.refill_min = o32.block_size, o32.blocks_per_cu, o32.refill_min
        oe.escalate_disk_edge = 0.02
        mse, ste = time_frames(acc, cam, metric, oe, args.steps, args.warmup)
        result["with_edge_escalation"] = {
            "value": n_rays / mse / 1e6, "ms_per_step": mse * 1e3, "rays_escalated": ste[-1]["rays_escalated"],
            "escalate_ms": float(np.mean([x["escalate_ms"] for x in ste])),
            "accuracy_vs_fp64": accuracy(acc.fetch_results(), ref, crit)}
        # the same frame with the critical band traced in FP64: rays within 0.5 M of the critical
        # curve go to the FP64 tracer on a second stream while the FP32 tracer runs the rest
        ob = default_opts(abi.PRECISION_FP32_ESCALATE)
        ob.block_size, ob.blocks_per_cu, ob.refill_min = o32.block_size, o32.blocks_per_cu, o32.refill_min
        ob.escalate_band = ob.critical_band
        msb, stb = time_frames(acc, cam, metric, ob, args.steps, args.warmup)
        result["with_band_escalation"] = {
            "value": n_rays / msb / 1e6, "ms_per_step": msb * 1e3, "escalate_band_m": ob.escalate_band,
            "rays_escalated": stb[-1]["rays_escalated"],
            "escalate_ms": float(np.mean([x["escalate_ms"] for x in stb])),
            "accuracy_vs_fp64": accuracy(acc.fetch_results(), ref, crit)}
    if not args.no_extra:
        result["other_workloads"] = other_workloads(acc, args.steps, args.warmup)
    if not args.no_surrogate and ref is not None:
        result["surrogate"] = surrogate_ablation(acc, cam, metric, ref, args.steps, args.warmup)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(cpu_threads())
        if "other_workloads" in result:
            result["cpu_baseline"]["image_writers"] = reference_writers(
                acc, result["other_workloads"]["image_writers_1080p"])
    acc.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
