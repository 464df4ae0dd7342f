"""Image writers: CPU restatement of the reference's per-pixel conversions and decoders of the
files. TEST INFRASTRUCTURE ONLY (tests/, bench.py's CPU baseline); the product never imports it.

Restated (numpy, every float step a separate float32 operation in the reference's order, pow in
FP64 rounded once to float):
  srgb8      PNGWriter::toSRGB8        Sirius.Render/Output/OUPN001A.h:103-112 (and
             ImageBuffer::toSRGB8, OUIB001A.h, the same steps)
  alpha8     PNGWriter::writeRGBA's alpha   OUPN001A.cpp:61-62
  aces8      BufferWriter::writePPM    Sirius.Render/Buffer/BFIO001A.h:29-36,44-68
  half_bits  tinyexr float_to_half_full (lib/tinyexr/tinyexr.h:985-1023), which
             EXRWriter::writeEXR requests (OUEW001A.cpp:155-157)
Decoders: PNG (zlib + the PNG filters, the row reconstruction in oracle_image.c), OpenEXR
scanline files (uncompressed, ZIPS, ZIP; HALF and FLOAT channels), PPM (P6), PFM, and the raw
float dump of BufferWriter::writeRaw (BFIO001A.h:71-86).

Pins (tests/test_images.py): every restatement reproduces the reference's own files for the
frames of tests/golden/ref_images.npz (written by the reference writers compiled in place, stb
and tinyexr), bit for bit; the PNG decoder agrees with stb_image and the EXR decoder with
tinyexr's loader on those files; half_bits agrees with IEEE round-to-nearest-even (numpy's
float16) wherever the two rules coincide, and rounds ties away from zero where they differ.
"""
import ctypes as C
import struct
import zlib

import numpy as np

F = np.float32


def _powf(x, y):
    return (x.astype(np.float64) ** np.float64(y)).astype(F)


def _to_u8(v):
    v = np.where(np.isnan(v), F(0), v)  # the C++ cast of NaN is undefined (x86 gives 0)
    return np.trunc(v).astype(np.uint8)


def srgb8(x):
    c = np.clip(np.asarray(x, F), F(0), F(1))
    c = np.where(np.isnan(x), np.asarray(x, F), c)  # std::clamp passes NaN through
    s = np.where(c <= F(0.0031308), F(12.92) * c, F(1.055) * _powf(c, F(1.0) / F(2.4)) - F(0.055))
    return _to_u8(np.clip(s * F(255.0) + F(0.5), F(0), F(255)))


def alpha8(a):
    a = np.asarray(a, F)
    return _to_u8(np.clip(a * F(255.0) + F(0.5), F(0), F(255)))


def aces8(x, exposure=1.0):
    x = np.asarray(x, F) * F(exposure)
    a, b, c, d, e = F(2.51), F(0.03), F(2.43), F(0.59), F(0.14)
    with np.errstate(all="ignore"):
        t = (x * (a * x + b)) / (x * (c * x + d) + e)
    t = np.where(np.isnan(t), t, np.clip(t, F(0), F(1)))
    t = _powf(t, F(1.0) / F(2.2))
    return _to_u8(np.where(np.isnan(t), t, np.clip(t * F(255.0), F(0), F(255))))


def half_bits(x):
    """float32 -> half bits as tinyexr converts: float denormals to zero, NaN to the quiet NaN
    0x7e00, overflow to infinity, otherwise the dropped bits rounded up when the first of them
    is set (ties away from zero)."""
    u = np.asarray(x, F).view(np.uint32).astype(np.int64)
    sign = (u >> 16) & 0x8000
    exp = (u >> 23) & 0xFF
    mant = u & 0x7FFFFF
    e = exp - 127 + 15
    normal = ((e << 10) | (mant >> 13)) + ((mant >> 12) & 1)
    shift = np.clip(14 - e, 1, 63)
    m = mant | 0x800000
    sub = np.where(14 - e <= 24, (m >> shift) + ((m >> (shift - 1)) & 1), 0)
    h = np.where(exp == 0, 0,
        np.where(exp == 255, 0x7C00 | np.where(mant != 0, 0x200, 0),
        np.where(e >= 31, 0x7C00,
        np.where(e <= 0, sub, normal))))
    return ((sign | h) & 0xFFFF).astype(np.uint16)


def half_to_float(h):
    return np.asarray(h, np.uint16).view(np.float16).astype(F)


# ---- decoders ---------------------------------------------------------------------------------
def png_decode(data):
    """(h, w, channels) uint8 of an 8-bit, non-interlaced grey/RGB/RGBA PNG; chunk CRCs checked."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n", "PNG signature"
    pos, idat, ihdr = 8, [], None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(typ + body) == crc, f"CRC of {typ}"
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat.append(body)
        elif typ == b"IEND":
            break
        pos += 12 + n
    w, h, depth, ctype, comp, filt, interlace = ihdr
    assert depth == 8 and comp == 0 and filt == 0 and interlace == 0, ihdr
    ch = {0: 1, 2: 3, 6: 4}[ctype]
    raw = zlib.decompress(b"".join(idat))
    assert len(raw) == h * (1 + w * ch), "filtered data size"
    from oracle import oracle_py as O
    L = O.lib()
    L.oracle_png_unfilter.restype = C.c_int
    L.oracle_png_unfilter.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_int, C.c_void_p]
    out = np.zeros((h, w, ch), np.uint8)
    assert L.oracle_png_unfilter(raw, h, w, ch, out.ctypes.data) == 0, "PNG filter type"
    return out


def png_filters(data):
    """The filter type byte of each row (for checking that the writer uses them all)."""
    pos, idat = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        if typ == b"IHDR":
            w, h = struct.unpack(">II", data[pos + 8:pos + 16])
            ch = {0: 1, 2: 3, 6: 4}[data[pos + 17]]
        elif typ == b"IDAT":
            idat.append(data[pos + 8:pos + 8 + n])
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8).reshape(h, 1 + w * ch)
    return raw[:, 0].copy()


def _exr_attrs(data):
    assert data[:4] == b"\x76\x2f\x31\x01", "EXR magic"
    version, = struct.unpack("<I", data[4:8])
    assert version & 0xFF == 2 and not version & 0x1E00, "single-part scanline EXR"
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        end2 = data.index(b"\0", end + 1)
        typ = data[end + 1:end2].decode()
        size, = struct.unpack("<i", data[end2 + 1:end2 + 5])
        attrs[name] = (typ, data[end2 + 5:end2 + 5 + size])
        pos = end2 + 5 + size
    return attrs, pos + 1


def _chlist(value):
    out, pos = [], 0
    while value[pos] != 0:
        end = value.index(b"\0", pos)
        name = value[pos:end].decode()
        ptype, = struct.unpack("<i", value[end + 1:end + 5])
        xs, ys = struct.unpack("<ii", value[end + 9:end + 17])
        out.append((name, ptype, xs, ys))
        pos = end + 17
    return out


def exr_decode(data):
    """{'attrs': {name: (type, bytes)}, 'channels': {name: array (h, w)}} of a scanline EXR;
    HALF channels come back as uint16 bits, FLOAT as float32."""
    attrs, pos = _exr_attrs(data)
    chans = _chlist(attrs["channels"][1])
    comp = attrs["compression"][1][0]
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines_per = {0: 1, 2: 1, 3: 16}[comp]
    nb = (h + lines_per - 1) // lines_per
    offsets = struct.unpack(f"<{nb}Q", data[pos:pos + 8 * nb])
    size_of = {1: 2, 2: 4}
    out = {name: np.zeros((h, w), np.uint16 if pt == 1 else F) for name, pt, _, _ in chans}
    for b, off in enumerate(offsets):
        y, n = struct.unpack("<ii", data[off:off + 8])
        blk = data[off + 8:off + 8 + n]
        lines = min(lines_per, y1 + 1 - y)
        raw_size = lines * w * sum(size_of[pt] for _, pt, _, _ in chans)
        if comp != 0 and n < raw_size:
            t = np.frombuffer(zlib.decompress(blk), np.uint8)
            # predictor: r[0] = t[0], r[j] = t[j] + r[j-1] - 128 (mod 256)
            r = (np.cumsum(np.concatenate([[np.int64(t[0])], t[1:].astype(np.int64) - 128])) % 256).astype(np.uint8)
            half = (len(r) + 1) // 2
            raw = np.empty(len(r), np.uint8)
            raw[0::2] = r[:half]
            raw[1::2] = r[half:]
        else:
            raw = np.frombuffer(blk, np.uint8)
        assert len(raw) == raw_size, "block size"
        p = 0
        for li in range(lines):
            for name, pt, _, _ in chans:
                k = size_of[pt] * w
                dt = np.uint16 if pt == 1 else F
                out[name][y - y0 + li] = np.frombuffer(raw[p:p + k].tobytes(), dt)
                p += k
    return {"attrs": attrs, "channels": out, "width": w, "height": h}


def ppm_decode(data):
    """(h, w, 3) uint8 of a binary PPM whose header is 'P6\\n<w> <h>\\n255\\n'."""
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"255", parts[:3]
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(h, w, 3)


def pfm_decode(data):
    """(h, w, 3) float32, top row first, of a little-endian colour PFM."""
    parts = data.split(b"\n", 3)
    assert parts[0] == b"PF" and float(parts[2]) < 0, parts[:3]
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], "<f4").reshape(h, w, 3)[::-1]


def raw_decode(data):
    w, h, c = struct.unpack("<iii", data[:12])
    return np.frombuffer(data[12:], "<f4").reshape(h, w, c)
