"""Writes the golden fixtures in tests/golden/ from the reference's own CPU code
(oracle/_ref/libsirius_ref.so, built by `make -C oracle ref` from the reference sources in
place) and from the FP64 oracle (the build's "pure FP64" columns).

Run in the build container (it needs the reference tree): python tests/golden/make_golden.py
(`python tests/golden/make_golden.py starfield`, `... images` or `... shading` writes only that
fixture).
Every fixture records its seed. The reference's GeodesicTracer::trace cannot be used as-is
(it ends a ray on its first rejected step, GTRC001A.cpp:269-277), so the trace fixtures use the
reference's Geodesic::integrateStepRK45 inside the harness loop of oracle/ref_harness.cpp.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle_py as O  # noqa: E402
from oracle import ref_py as RF  # noqa: E402
from sirius_amd import abi  # noqa: E402


def metric_points(seed=0x51121001):
    rng = np.random.default_rng(seed)
    ks_params, ks_x, ks_g, ks_dg = [], [], [], []
    while len(ks_x) < 256:
        a = rng.uniform(-0.99, 0.99)
        Q = rng.uniform(0, 0.5) * np.sqrt(max(1 - a * a, 0))
        L = rng.choice([0.0, rng.uniform(0, 1e-3)])
        x = np.r_[rng.uniform(-5, 5), rng.uniform(-25, 25, 3)]
        if np.linalg.norm(x[1:]) < 2.5:
            continue
        g, dg = RF.ks_evaluate(1.0, a, Q, L, x)
        ks_params.append((1.0, a, Q, L)); ks_x.append(x); ks_g.append(g); ks_dg.append(dg)
    mt_x, mt_g, mt_dg = [], [], []
    for _ in range(64):
        x = np.r_[0.0, rng.uniform(1.1, 20.0), rng.uniform(0.1, np.pi - 0.1), rng.uniform(0, 2 * np.pi)]
        g, dg = RF.mt_evaluate(1.0, x)
        mt_x.append(x); mt_g.append(g); mt_dg.append(dg)
    wp_x, wp_g, wp_dg = [], [], []
    for _ in range(64):
        x = np.r_[0.0, rng.uniform(-3, 3, 3)]
        g, dg = RF.warp_evaluate(0.5, 1.0, 8.0, x)
        wp_x.append(x); wp_g.append(g); wp_dg.append(dg)
    np.savez_compressed(os.path.join(HERE, "ref_metric.npz"), seed=seed,
                        ks_params=np.array(ks_params), ks_x=np.array(ks_x), ks_g=np.array(ks_g),
                        ks_dg=np.array(ks_dg), mt_b0=1.0, mt_x=np.array(mt_x), mt_g=np.array(mt_g),
                        mt_dg=np.array(mt_dg), warp_params=np.array([0.5, 1.0, 8.0]),
                        warp_x=np.array(wp_x), warp_g=np.array(wp_g), warp_dg=np.array(wp_dg))


def camera_and_closed_forms():
    cams = [(30.0, np.deg2rad(80), 0.0, 30.0, 256, 256, 0.0),
            (50.0, np.deg2rad(75), 0.0, 20.0, 1920, 1080, 0.9)]
    rows = []
    for r, th, ph, fov, w, h, a in cams:
        for px, py in [(0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1), (w // 2, h // 2), (w // 3, 2 * h // 3),
                       (17, 5), (w - 40, h // 7)]:
            o, d = RF.camera_ray(r, th, ph, fov, w, h, px, py)
            x, k = RF.init_state(1.0, a, 0.0, 0.0, r, th, ph, fov, w, h, px, py)
            rows.append(np.r_[r, th, ph, fov, w, h, a, px, py, o, d, x, k])
    spins = np.array([0.0, 0.1, 0.5, 0.9, 0.99, -0.5, -0.9])
    isco = np.array([RF.isco(a) for a in spins])
    circ_r, circ = [], []
    for a in (0.0, 0.5, 0.9):
        for r in (7.0, 10.0, 25.0):
            circ_r.append((a, r)); circ.append(RF.circular(a, r))
    ptq = np.array([[a, r, RF.pt_q(a, r)] for a in (0.0, 0.9) for r in (9.0, 15.0, 30.0)])
    np.savez_compressed(os.path.join(HERE, "ref_camera_closed.npz"), camera_rows=np.array(rows),
                        spins=spins, isco=isco, circular_ar=np.array(circ_r), circular_E_L_Om=np.array(circ),
                        pt_q=ptq)


def trace_fixture(name, M, a, r_obs, incl, fov, w, h, disk, n_bins, pix, seed):
    R = RF.trace(M, a, 0.0, 0.0, r_obs, np.deg2rad(incl), 0.0, fov, w, h, pix, r_cap=1.01 * (1 + np.sqrt(1 - a * a)),
                 r_esc=1000.0, disk=(disk[0] if disk[0] > 0 else RF.isco(a), disk[1]))
    cam = abi.camera(r_obs, incl, fov, w, h, disk=disk)
    m = abi.metric_ks(M, a)
    opts = O.default_opts(abi.PRECISION_FP64)
    opts.n_bins = n_bins
    out, _ = O.trace([cam], [m], opts, ray_ids=pix, with_state=True)
    crit = np.array([O.critical_distance(cam, m, int(i % w), int(i // w)) for i in pix])
    np.savez_compressed(os.path.join(HERE, name), seed=seed, pix=pix, width=w, height=h, a=a, r_obs=r_obs,
                        incl_deg=incl, fov=fov, disk=np.array(disk), n_bins=n_bins,
                        ref_cls=R["cls"].astype(np.int8), ref_steps=R["steps"], ref_x=R["x"], ref_k=R["k"],
                        ref_disk_r=R["disk_r"], ref_lam=R["lam"],
                        fp64_cls=out.cls, fp64_theta=out.theta, fp64_phi=out.phi, fp64_lam=out.lam,
                        fp64_disk_r=out.disk_r, fp64_g=out.redshift, fp64_I=out.intensity,
                        fp64_steps=out.steps, fp64_state=out.state, crit=crit)


def starfield_fixture(seed=0x51121008):
    """The reference's star catalogue generator and sampler: two catalogues (the two celestial
    spheres of the wormhole scene: 4096 candidates, 1 to 1000 pc, magnitude limit 20, seeds 42
    and 43; and the reference's default limit with a clamped magnitude limit) and the sampled
    star light at directions exactly on stars, at offsets of 1e-5 to 1e-2 rad from stars, and
    at random directions, with and without the depth-of-field blur."""
    rng = np.random.default_rng(seed)
    cats = {"near": RF.starfield_generate(4096, 1.0, 1000.0, 20.0, 42),
            "far": RF.starfield_generate(4096, 1.0, 1000.0, 20.0, 43),
            "clamped": RF.starfield_generate(2000, 1.0, 10000.0, 25.0, 7)}
    out = {"seed": seed}
    for name, cat in cats.items():
        out[f"cat_{name}"] = cat
    cat = cats["near"]
    on = cat[:64, :3]
    off = []
    for ang in (1e-5, 1e-4, 1e-3, 1e-2):
        v = cat[64:96, :3].astype(np.float64)
        t = rng.normal(size=v.shape)
        t -= (t * v).sum(1, keepdims=True) * v
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        d = np.cos(ang) * v + np.sin(ang) * t
        off.append(d / np.linalg.norm(d, axis=1, keepdims=True))
    rnd = rng.normal(size=(256, 3))
    rnd /= np.linalg.norm(rnd, axis=1, keepdims=True)
    dirs = np.concatenate([on, np.concatenate(off), rnd]).astype(np.float32)
    out["dirs"] = dirs
    out["rgb_dof"] = RF.starfield_sample(50.0, 100.0, 1.0, 1, cat, dirs)
    out["rgb_nodof"] = RF.starfield_sample(0.0, 100.0, 2.0, 0, cat, dirs)
    out["rgb_far_dof"] = RF.starfield_sample(50.0, 100.0, 1.0, 1, cats["far"], dirs)
    np.savez_compressed(os.path.join(HERE, "ref_starfield.npz"), **out)


def image_frames(seed=0x51121009):
    """Three float RGBA frames for the image writers. ldr (35 x 37): values across [-0.2, 1.3],
    around the sRGB threshold 0.0031308 (it, and the floats next to it), small values down to
    1e-6 and HDR values up to 50 for the ACES curve, alpha over [-0.1, 1.1]. hdr (35 x 37): NaN,
    +-Inf, +-0, float denormals, half denormals, exact half ties, the half overflow boundary
    (65504, 65519, 65520) and log-uniform values of both signs from 1e-9 to 1e6. noise (20 x 33):
    random finite floats of random exponents, which deflate does not shrink (tinyexr then stores
    the block plain)."""
    rng = np.random.default_rng(seed)
    h, w = 35, 37
    ldr = rng.uniform(-0.2, 1.3, (h, w, 4)).astype(np.float32)
    ldr[..., 3] = rng.uniform(-0.1, 1.1, (h, w))
    t = np.float32(0.0031308)
    special = [t, np.nextafter(t, np.float32(0)), np.nextafter(t, np.float32(1)), 0.0, 1.0, 0.5, 1e-6, 1e-4, 1e-3,
               0.04045, 0.99999994, 1.0000001]
    flat = ldr.reshape(-1, 4)
    flat[:len(special), :3] = np.array(special, np.float32)[:, None]
    flat[40:140, :3] = np.exp(rng.uniform(np.log(1e-6), np.log(1e-2), (100, 3))).astype(np.float32)
    flat[200:300, :3] = rng.uniform(1.0, 50.0, (100, 3)).astype(np.float32)
    hdr = (np.exp(rng.uniform(np.log(1e-9), np.log(1e6), (h, w, 4))) * rng.choice([-1.0, 1.0], (h, w, 4))).astype(np.float32)
    ties = [1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 2.0 ** -15 * (1 + 2.0 ** -11), 2.0 ** -20, 3 * 2.0 ** -25,
            2.0 ** -25, 2.0 ** -24, 1.5 * 2.0 ** -24, 65504.0, 65519.0, 65520.0, 1e5, 6.1e-5, 6.0e-8]
    specials = [np.nan, np.inf, -np.inf, 0.0, -0.0, 1e-40, -1e-41] + ties + [-v for v in ties]
    fh = hdr.reshape(-1, 4)
    fh[:len(specials), :] = np.array(specials, np.float32)[:, None]
    fh[len(specials):2 * len(specials), :] = np.roll(np.array(specials, np.float32), 1)[:, None]
    noise = rng.integers(0, 2 ** 32, (20, 33, 4), dtype=np.uint64).astype(np.uint32)
    noise = noise & np.uint32(0x8FFFFFFF) | np.uint32(0x30000000)  # finite, exponents spread
    return {"ldr": ldr, "hdr": hdr, "noise": noise.view(np.float32)}


def images_fixture():
    """The reference writers' files for the frames of image_frames(): PNG (PNGWriter::writeRGBA)
    and the two PPMs (BufferWriter::writePPM at exposure 1 and 1.7, EXRWriter::writePPM) of the
    ldr frame; EXR (EXRWriter::writeEXR), PFM (EXRWriter::writePFM) and the raw dump
    (BufferWriter::writeRaw) of the hdr frame; EXR of the noise frame; and the decodings by the
    reference's vendored stb_image and tinyexr."""
    fr = image_frames()
    out = {f"frame_{k}": v for k, v in fr.items()}
    b = lambda data: np.frombuffer(data, np.uint8)
    ldr, hdr = fr["ldr"], fr["hdr"]
    png = RF.write_png(ldr)
    out["png_ldr"] = b(png)
    out["png_ldr_stb"] = RF.load_png(png, *ldr.shape[:2])
    out["ppm_aces_ldr_e1"] = b(RF.write_ppm_aces(ldr, 1.0))
    out["ppm_aces_ldr_e1.7"] = b(RF.write_ppm_aces(ldr, 1.7))
    out["ppm_srgb_ldr"] = b(RF.write_ppm_srgb(ldr[..., :3]))
    for name in ("hdr", "noise"):
        exr = RF.write_exr(fr[name])
        out[f"exr_{name}"] = b(exr)
        out[f"exr_{name}_tinyexr"] = RF.load_exr(exr, *fr[name].shape[:2])
    out["pfm_hdr"] = b(RF.write_pfm(hdr[..., :3]))
    out["raw_hdr"] = b(RF.write_raw(hdr))
    np.savez_compressed(os.path.join(HERE, "ref_images.npz"), **out)


def shading_fixture(seed=0x5112100C):
    """The reference's disk-colour functions: blackbodyToRGB over 1000 K - 1e5 K,
    applyColorMode(TrueColor) on random (T_emit, g, intensity) and applyLimbDarkening on random
    colours and cosines (endpoints included)."""
    rng = np.random.default_rng(seed)
    T = np.concatenate([np.geomspace(1000.0, 1e5, 400), rng.uniform(1000.0, 1e5, 200), [5778.0, 6500.0]])
    te = rng.uniform(0.02, 4.0, 300).astype(np.float32)
    g = rng.uniform(0.1, 2.0, 300).astype(np.float32)
    inten = rng.uniform(0.0, 50.0, 300).astype(np.float32)
    rgb = rng.uniform(0.0, 2.0, (300, 3)).astype(np.float32)
    cos_i = np.concatenate([rng.uniform(-0.2, 1.0, 296), [0.0, 1.0, -1e-3, 0.5]])
    np.savez_compressed(os.path.join(HERE, "ref_shading.npz"), seed=seed, T=T, bb_rgb=RF.blackbody_rgb(T),
                        tc_T=te, tc_g=g, tc_I=inten, tc_rgb=RF.true_color(te, g, inten), limb_in=rgb, limb_cos=cos_i,
                        limb_rgb=RF.limb_darkening(rgb, cos_i))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "shading":
        shading_fixture()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "starfield":
        starfield_fixture()
        return
    if len(sys.argv) > 1 and sys.argv[1] == "images":
        images_fixture()
        return
    if not RF.available():
        sys.exit("oracle/_ref/libsirius_ref.so missing: run make -C oracle ref (needs the reference tree)")
    metric_points()
    camera_and_closed_forms()
    # C1: Schwarzschild, r=30, 80 deg, FOV 30, 256^2, disk [6, 20], 8 bins: every 4th pixel per axis
    yy, xx = np.meshgrid(np.arange(0, 256, 4), np.arange(0, 256, 4), indexing="ij")
    trace_fixture("ref_trace_c1.npz", 1.0, 0.0, 30.0, 80.0, 30.0, 256, 256, (6.0, 20.0), 8,
                  (yy * 256 + xx).ravel().astype(np.int64), seed=0)
    # C2: Kerr a=0.9, r=50, 75 deg, FOV 20, 1920x1080, disk [ISCO, 30], 16 bins: 3000 random pixels
    seed = 0x51121002
    pix = np.sort(np.random.default_rng(seed).choice(1920 * 1080, 3000, replace=False)).astype(np.int64)
    trace_fixture("ref_trace_c2.npz", 1.0, 0.9, 50.0, 75.0, 20.0, 1920, 1080, (0.0, 30.0), 16, pix, seed=seed)
    starfield_fixture()
    images_fixture()
    shading_fixture()


if __name__ == "__main__":
    main()
