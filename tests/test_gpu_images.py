"""Image writers and the ray result dump on the device, through the C-ABI.

1. Each format encoded by the library from the fixture frames decodes to exactly what the
   reference writer's file decodes to (tests/golden/ref_images.npz: PNGWriter, EXRWriter,
   BufferWriter); the uncompressed formats (PPM, PFM, raw) are byte-identical files.
2. At full size (a 1920 x 1080 frame of values across the sRGB, ACES and half ranges) the files
   decode to the oracle's restatement of the reference conversions, whatever the thread count
   or zlib level; the framebuffer path and a device array give the same files as the host copy.
3. The ray dump of a trace reads back bit for bit, with the records that made it, and rays
   re-traced by the oracle from those records alone agree with it.
"""
import numpy as np
import pytest

from oracle import oracle_image as I
from oracle import oracle_py as O
from sirius_amd import abi
from sirius_amd.accelerator import SiriusHipError
from sirius_amd.raydump import read_rays

pytestmark = pytest.mark.gpu


def _g(golden):
    return golden("ref_images.npz")


def test_png_matches_reference(acc, golden):
    g = _g(golden)
    print("[case] png 37x35", flush=True)
    data, st = acc.encode_image(abi.IMAGE_PNG, g["frame_ldr"])
    assert np.array_equal(I.png_decode(data), I.png_decode(g["png_ldr"].tobytes()))
    assert st.file_bytes == len(data) and st.payload_bytes == 35 * (1 + 4 * 37)


@pytest.mark.parametrize("key,exposure", [("ppm_aces_ldr_e1", 1.0), ("ppm_aces_ldr_e1.7", 1.7)])
def test_ppm_aces_is_the_reference_file(acc, golden, key, exposure):
    g = _g(golden)
    print(f"[case] ppm aces exposure {exposure}", flush=True)
    data, _ = acc.encode_image(abi.IMAGE_PPM, g["frame_ldr"], exposure=exposure)
    assert data == g[key].tobytes()


def test_ppm_srgb_pfm_raw_are_the_reference_files(acc, golden):
    g = _g(golden)
    print("[case] ppm srgb / pfm / raw", flush=True)
    assert acc.encode_image(abi.IMAGE_PPM_SRGB, g["frame_ldr"])[0] == g["ppm_srgb_ldr"].tobytes()
    assert acc.encode_image(abi.IMAGE_PFM, g["frame_hdr"])[0] == g["pfm_hdr"].tobytes()
    assert acc.encode_image(abi.IMAGE_RAW, g["frame_hdr"])[0] == g["raw_hdr"].tobytes()


@pytest.mark.parametrize("name", ["hdr", "noise"])
def test_exr_matches_reference(acc, golden, name):
    g = _g(golden)
    print(f"[case] exr {name}", flush=True)
    data, _ = acc.encode_image(abi.IMAGE_EXR, g[f"frame_{name}"])
    got, ref = I.exr_decode(data), I.exr_decode(g[f"exr_{name}"].tobytes())
    assert got["attrs"] == ref["attrs"]
    for c in "ABGR":
        assert np.array_equal(got["channels"][c], ref["channels"][c]), c


def _big_frame(h=1080, w=1920, seed=11):
    rng = np.random.default_rng(seed)
    f = np.empty((h, w, 4), np.float32)
    # smooth image-like regions (compressible) and noisy ones, over the ranges the curves care about
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    f[..., 0] = (xx / w) * 1.3 - 0.1
    f[..., 1] = np.exp(np.log(1e-6) * (yy / h)).astype(np.float32)
    f[..., 2] = rng.uniform(-1.0, 60.0, (h, w))
    f[..., 3] = rng.uniform(-0.1, 1.1, (h, w))
    f[h // 2:h // 2 + 100] = rng.uniform(-7e4, 7e4, (100, w, 4))
    return f


def test_full_frame_formats_decode_to_the_restatement(acc):
    f = _big_frame()
    print("[case] 1920x1080 png / exr / ppm", flush=True)
    png, st = acc.encode_image(abi.IMAGE_PNG, f)
    want = np.concatenate([I.srgb8(f[..., :3]), I.alpha8(f[..., 3:])], axis=-1)
    assert np.array_equal(I.png_decode(png), want)
    filters = I.png_filters(png)
    assert len(set(filters.tolist())) >= 2  # the per-row choice is made, not a fixed filter
    exr, _ = acc.encode_image(abi.IMAGE_EXR, f)
    d = I.exr_decode(exr)
    for k, c in enumerate("RGBA"):
        assert np.array_equal(d["channels"][c], I.half_bits(f[..., k])), c
    ppm, _ = acc.encode_image(abi.IMAGE_PPM, f, exposure=0.8)
    assert np.array_equal(I.ppm_decode(ppm), I.aces8(f[..., :3], 0.8))
    # one thread or many, any level: the same pixels
    for threads, level in ((1, 1), (3, 9), (16, 0)):
        png2, _ = acc.encode_image(abi.IMAGE_PNG, f, zlib_level=level, threads=threads)
        assert np.array_equal(I.png_decode(png2), want), (threads, level)
        exr2, _ = acc.encode_image(abi.IMAGE_EXR, f, zlib_level=level, threads=threads)
        assert np.array_equal(I.exr_decode(exr2)["channels"]["G"], d["channels"]["G"]), (threads, level)


def test_framebuffer_and_device_array_paths(acc, tmp_path):
    from devmem import DeviceArray
    w, h = 96, 54
    acc.initialise(w, h)
    acc.uploadBackground(np.zeros((1, 1, 4), np.uint8))
    m = abi.metric_ks(1.0, 0.9)
    cam = abi.camera(50.0, 75.0, 20.0, w, h, disk=(0.0, 30.0))
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    print("[case] framebuffer -> png / exr", flush=True)
    acc.launch(cam, m, o)
    fb = acc.getFrameBuffer().reshape(h, w, 4).copy()
    for fmt in (abi.IMAGE_PNG, abi.IMAGE_EXR, abi.IMAGE_PPM, abi.IMAGE_RAW):
        from_fb, _ = acc.encode_image(fmt)
        from_host, _ = acc.encode_image(fmt, fb)
        assert from_fb == from_host, fmt
        dev = DeviceArray(fb)
        from_dev, _ = acc.encode_image(fmt, dev)
        assert from_dev == from_host, fmt
    p = tmp_path / "frame.exr"
    st = acc.write_image(str(p), abi.IMAGE_EXR)
    assert p.read_bytes() == acc.encode_image(abi.IMAGE_EXR)[0] and st.file_bytes == p.stat().st_size


def test_image_errors(acc, tmp_path):
    with pytest.raises(SiriusHipError) as e:
        acc.encode_image(abi.IMAGE_PNG, np.zeros((0, 4, 4), np.float32))
    assert e.value.rc == abi.E_INVALID
    with pytest.raises(SiriusHipError) as e:
        acc.encode_image(99, np.zeros((2, 2, 4), np.float32))
    assert e.value.rc == abi.E_INVALID
    with pytest.raises(SiriusHipError) as e:
        acc.write_image(str(tmp_path / "no" / "such" / "dir.png"), abi.IMAGE_PNG, np.zeros((2, 2, 4), np.float32))
    assert e.value.rc == abi.E_IO


def test_ray_dump_round_trip_and_audit(acc, tmp_path):
    m = abi.metric_ks(1.0, 0.9)
    cams = [abi.camera(50.0, 75.0, 20.0, 64, 36, disk=(0.0, 30.0)),
            abi.camera(30.0, 60.0, 30.0, 32, 32, disk=(0.0, 20.0))]
    o = O.default_opts(abi.PRECISION_FP32_ESCALATE)
    print("[case] ray dump", flush=True)
    rays, _ = acc.trace_batch(cams, [m], o, with_state=True)
    p = tmp_path / "rays.bin"
    acc.write_rays(str(p), cams, [m], o)
    d = read_rays(str(p))
    assert d["n_rays"] == rays.n and d["n_bins"] == o.n_bins
    for name, attr in (("cls", "cls"), ("flags", "flags"), ("theta", "theta"), ("phi", "phi"),
                       ("lambda", "lam"), ("disk_r", "disk_r"), ("redshift", "redshift"),
                       ("null_drift", "null_drift"), ("steps", "steps"), ("intensity", "intensity"),
                       ("state", "state")):
        assert np.array_equal(d[name], getattr(rays, attr), equal_nan=True), name
    assert [bytes(c) for c in d["cameras"]] == [bytes(c) for c in cams]
    assert bytes(d["opts"]) == bytes(o)
    # a field subset leaves the others out
    acc.write_rays(str(p), cams, [m], None, fields=1 << abi.RAYS_FIELDS.index("theta"))
    d2 = read_rays(str(p))
    assert set(abi.RAYS_FIELDS) & set(d2) == {"theta"} and d2["opts"] is None
    # the audit: rays re-traced by the oracle from the dump's own records agree with it
    rng = np.random.default_rng(1)
    ids = np.sort(rng.choice(d["n_rays"], 64, replace=False))
    ref, _ = O.trace(d["cameras"], d["metrics"], O.default_opts(abi.PRECISION_FP64), ray_ids=ids)
    agree = ref.cls == d["cls"][ids]
    assert agree.mean() >= 0.95
    esc = agree & (ref.cls == abi.RAY_ESCAPED)
    if esc.any():
        dth = np.abs(ref.theta[esc] - d["theta"][ids][esc])
        assert np.median(dth) < 1e-4


def test_ray_dump_needs_results(tmp_path):
    from sirius_amd.accelerator import HipAccelerator
    a = HipAccelerator()
    try:
        with pytest.raises(SiriusHipError) as e:
            a.write_rays(str(tmp_path / "x.bin"))
        assert e.value.rc == abi.E_STATE
        with pytest.raises(SiriusHipError) as e:
            a.encode_image(abi.IMAGE_PNG)
        assert e.value.rc == abi.E_STATE
    finally:
        a.close()


@pytest.mark.parametrize("h,w", [(1, 1), (1, 17), (17, 1), (16, 3), (33, 5)])
def test_edge_sizes_decode_to_the_restatement(acc, h, w):
    """Frames of one pixel, one row or column, and EXR heights on and across the 16-line block
    boundary (PNGWriterTest.GammaCorrection, TSOF009A.cpp:155-178, writes a 1x1 frame of 0.5)."""
    rng = np.random.default_rng(h * 100 + w)
    f = rng.uniform(-0.2, 1.3, (h, w, 4)).astype(np.float32)
    if (h, w) == (1, 1):
        f[:] = 0.5
    print(f"[case] edge size {h}x{w}", flush=True)
    png, _ = acc.encode_image(abi.IMAGE_PNG, f)
    want = np.concatenate([I.srgb8(f[..., :3]), I.alpha8(f[..., 3:])], axis=-1)
    assert np.array_equal(I.png_decode(png), want)
    if (h, w) == (1, 1):
        assert list(I.png_decode(png)[0, 0, :3]) == [188, 188, 188] and len(png) < 1000
    d = I.exr_decode(acc.encode_image(abi.IMAGE_EXR, f)[0])
    for k, c in enumerate("RGBA"):
        assert np.array_equal(d["channels"][c], I.half_bits(f[..., k])), c
    assert np.array_equal(I.ppm_decode(acc.encode_image(abi.IMAGE_PPM, f)[0]), I.aces8(f[..., :3]))
    assert np.array_equal(I.pfm_decode(acc.encode_image(abi.IMAGE_PFM, f)[0]), f[..., :3])
