"""Image writers, on the CPU: the oracle's restatements and decoders against the reference's own
files.

tests/golden/ref_images.npz holds the files the reference writers (PNGWriter::writeRGBA,
EXRWriter::writeEXR / writePPM / writePFM, BufferWriter::writePPM / writeRaw;
src/Sirius.Render/Output/OUPN001A.*, OUEW001A.*, src/Sirius.Render/Buffer/BFIO001A.h, with the
stb_image_write, tinyexr and miniz sources vendored with the reference) wrote for three float
frames, and the decodings of the PNG and EXR files by the reference's own stb_image and tinyexr
(tests/golden/make_golden.py images).

Pins: the oracle's PNG and EXR decoders agree with stb_image and tinyexr; the restated
conversions (sRGB, alpha, ACES + gamma, tinyexr's float -> half) reproduce the reference files
bit for bit; the half conversion also agrees with IEEE round-to-nearest-even wherever the two
rules coincide.
"""
import os
import struct

import numpy as np
import pytest

from oracle import oracle_image as I

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_images.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLDEN)


def _bytes(g, key):
    return g[key].tobytes()


def test_png_decoder_matches_stb(g):
    assert np.array_equal(I.png_decode(_bytes(g, "png_ldr")), g["png_ldr_stb"])


@pytest.mark.parametrize("name", ["hdr", "noise"])
def test_exr_decoder_matches_tinyexr(g, name):
    d = I.exr_decode(_bytes(g, f"exr_{name}"))
    ref = g[f"exr_{name}_tinyexr"]
    for k, c in enumerate("RGBA"):
        got = I.half_to_float(d["channels"][c])
        assert np.array_equal(got, ref[..., k], equal_nan=True), c


def test_noise_frame_exercises_plain_blocks(g):
    """The noise frame's file has a compressed block (deflate gains a little on its first 16
    lines) and a block stored plain (the last 4 lines, which deflate does not shrink)."""
    data = _bytes(g, "exr_noise")
    attrs, pos = I._exr_attrs(data)
    off = struct.unpack("<2Q", data[pos:pos + 16])
    sizes = [struct.unpack("<ii", data[o:o + 8])[1] for o in off]
    assert sizes[0] < 16 * 33 * 8 and sizes[1] == 4 * 33 * 8


def test_srgb_restatement_reproduces_reference_png(g):
    f = g["frame_ldr"]
    want = np.concatenate([I.srgb8(f[..., :3]), I.alpha8(f[..., 3:])], axis=-1)
    assert np.array_equal(I.png_decode(_bytes(g, "png_ldr")), want)


@pytest.mark.parametrize("key,exposure", [("ppm_aces_ldr_e1", 1.0), ("ppm_aces_ldr_e1.7", 1.7)])
def test_aces_restatement_reproduces_reference_ppm(g, key, exposure):
    f = g["frame_ldr"]
    h, w = f.shape[:2]
    want = f"P6\n{w} {h}\n255\n".encode() + I.aces8(f[..., :3], exposure).tobytes()
    assert _bytes(g, key) == want


def test_srgb_ppm_restatement_reproduces_reference(g):
    f = g["frame_ldr"]
    h, w = f.shape[:2]
    assert _bytes(g, "ppm_srgb_ldr") == f"P6\n{w} {h}\n255\n".encode() + I.srgb8(f[..., :3]).tobytes()


@pytest.mark.parametrize("name", ["hdr", "noise"])
def test_half_restatement_reproduces_reference_exr(g, name):
    f = g[f"frame_{name}"]
    d = I.exr_decode(_bytes(g, f"exr_{name}"))
    assert [c for c in d["channels"]] == ["A", "B", "G", "R"]
    for k, c in enumerate("RGBA"):
        assert np.array_equal(d["channels"][c], I.half_bits(f[..., k])), c


def test_pfm_and_raw_layout_reproduce_reference(g):
    f = g["frame_hdr"]
    h, w = f.shape[:2]
    pfm = f"PF\n{w} {h}\n-1.0\n".encode() + np.ascontiguousarray(f[::-1, :, :3]).astype("<f4").tobytes()
    assert _bytes(g, "pfm_hdr") == pfm
    raw = struct.pack("<iii", w, h, 4) + f.astype("<f4").tobytes()
    assert _bytes(g, "raw_hdr") == raw
    assert np.array_equal(I.pfm_decode(pfm), f[..., :3], equal_nan=True)
    assert np.array_equal(I.raw_decode(raw), f, equal_nan=True)


def test_half_bits_closed_forms():
    rng = np.random.default_rng(3)
    x = (np.exp(rng.uniform(np.log(1e-9), np.log(7e4), 200000)) * rng.choice([-1, 1], 200000)).astype(np.float32)
    h = I.half_bits(x)
    with np.errstate(over="ignore"):
        rne = x.astype(np.float16).view(np.uint16)
    # the two rules differ only on exact ties (a set first dropped bit and nothing after it)
    u = x.view(np.uint32).astype(np.int64)
    e = ((u >> 23) & 0xFF) - 112
    drop = np.where(e > 0, 13, np.clip(14 - e, 1, 40))
    rem = (u & 0x7FFFFF | np.where(e > 0, 0, 0x800000)) & ((1 << drop) - 1)
    tie = rem == (1 << (drop - 1))
    assert np.array_equal(h[~tie], rne[~tie])
    f = np.float32
    cases = {1.0: 0x3C00, 65504.0: 0x7BFF, 65519.0: 0x7BFF, 65520.0: 0x7C00, 1e5: 0x7C00,
             1.0 + 2.0 ** -11: 0x3C01,          # tie: rounded away (IEEE would give 0x3C00)
             1.0 + 3 * 2.0 ** -11: 0x3C02,      # tie: away from zero, which is even here too
             2.0 ** -24: 0x0001, 2.0 ** -25: 0x0001, 2.0 ** -26: 0x0000, 1e-40: 0x0000,
             -0.0: 0x8000, -2.0: 0xC000, float("inf"): 0x7C00, float("-inf"): 0xFC00, float("nan"): 0x7E00}
    for v, want in cases.items():
        assert int(I.half_bits(np.array([v], f))[0]) == want, (v, hex(int(I.half_bits(np.array([v], f))[0])))


def test_srgb_and_aces_closed_forms():
    f = np.float32
    # the sRGB curve's anchors: 0 -> 0, 1 -> 255, the linear segment below 0.0031308, 0.5 -> 188
    assert list(I.srgb8(np.array([0.0, 1.0, 2.0, -1.0, 0.5], f))) == [0, 255, 255, 0, 188]
    assert int(I.srgb8(np.array([0.002], f))[0]) == int(f(12.92) * f(0.002) * f(255) + f(0.5))
    # ACES: 0 -> 0, saturates to 255 for large values; monotone
    v = np.linspace(0, 20, 2001, dtype=f)
    a = I.aces8(v).astype(int)
    assert a[0] == 0 and a[-1] == 255 and np.all(np.diff(a) >= 0)


def test_png_filters_round_trip():
    """The decoder undoes every PNG filter: rows filtered here with each type in turn."""
    import zlib
    rng = np.random.default_rng(5)
    h, w, bpp = 9, 13, 4
    img = rng.integers(0, 256, (h, w * bpp)).astype(np.int64)
    rows = []
    for y in range(h):
        t = y % 5
        x = img[y]
        a = np.concatenate([np.zeros(bpp, np.int64), x[:-bpp]])
        b = img[y - 1] if y else np.zeros_like(x)
        c = np.concatenate([np.zeros(bpp, np.int64), b[:-bpp]])
        p = a + b - c
        pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
        pae = np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))
        f = [x, x - a, x - b, x - (a + b) // 2, x - pae][t]
        rows.append(np.concatenate([[t], f % 256]).astype(np.uint8))
    raw = np.concatenate(rows).tobytes()

    def chunk(typ, body):
        return struct.pack(">I", len(body)) + typ + body + struct.pack(">I", zlib.crc32(typ + body))
    png = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0))
           + chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
    assert np.array_equal(I.png_decode(png).reshape(h, -1), img.astype(np.uint8))


def test_ray_dump_reader_follows_the_documented_layout(tmp_path):
    """A dump assembled byte by byte from the layout in include/sirius_hip.h reads back."""
    import ctypes as C
    from sirius_amd import abi
    from sirius_amd.raydump import read_rays
    n, nb = 5, 3
    cams = [abi.camera(50.0, 75.0, 20.0, 5, 1)]
    mets = [abi.metric_ks(a=0.9)]
    fields = (1 << 0) | (1 << 2) | (1 << 9) | (1 << 10)  # cls, theta, intensity, state
    head = bytearray(128)
    head[:8] = b"SIRRAYS1"
    struct.pack_into("<IIQIiIiII", head, 8, 128, nb, n, fields, 1, C.sizeof(abi.Camera), 1,
                     C.sizeof(abi.Metric), 0)
    body = bytes(head) + bytes(cams[0]) + bytes(mets[0])
    cls = np.arange(n, dtype=np.uint8)
    theta = np.linspace(0, 1, n, dtype=np.float32)
    inten = np.arange(nb * n, dtype=np.float32)
    state = np.arange(8 * n, dtype=np.float64)
    for arr in (cls, theta, inten, state):
        body += b"\0" * ((8 - len(body) % 8) % 8) + arr.tobytes()
    p = tmp_path / "rays.bin"
    p.write_bytes(body)
    d = read_rays(str(p))
    assert d["n_rays"] == n and d["n_bins"] == nb and d["opts"] is None
    assert np.array_equal(d["cls"], cls) and np.array_equal(d["theta"], theta)
    assert np.array_equal(d["intensity"], inten.reshape(nb, n)) and np.array_equal(d["state"], state.reshape(n, 8))
    assert "phi" not in d and bytes(d["cameras"][0]) == bytes(cams[0]) and d["metrics"][0].a == 0.9
