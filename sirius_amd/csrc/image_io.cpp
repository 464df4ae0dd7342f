// image_io.cpp — image writers and the ray result dump.
//
// A frame is converted on the device (image_kernels.hip) from the framebuffer or a caller's
// RGBA array, the converted payload is copied once into pinned host memory, and the host packs
// the file: PPM / PFM / raw headers, PNG chunks with one zlib stream, or an EXR scanline file
// with one ZIP block per 16 lines. Compression runs on a pool of host threads: EXR blocks are
// independent by the format; the PNG stream is deflated in pieces, each primed with the 32 KiB
// before it and ended on a byte boundary (a sync flush), which concatenate into one valid
// stream whose Adler-32 is combined from the pieces'.
//
// The files decode to the pixels of the reference writers (Sirius.Render/Output/OUPN001A.cpp,
// OUEW001A.*, Sirius.Render/Buffer/BFIO001A.h); compressed bytes differ (other deflate
// choices, other PNG row filters), decoded pixels do not.
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sirius_hip.h"
#include "sirius_internal.hpp"

namespace sirius {

hipError_t launch_ldr_pixels(int mode, const float* fb, int64_t n, float exposure, uint8_t* out, hipStream_t st);
hipError_t launch_png_filter(const uint8_t* px, int w, int h, int bpp, uint8_t* out, hipStream_t st);
hipError_t launch_exr_zip_prepass(const float* fb, int w, int h, uint8_t* out, hipStream_t st);
hipError_t launch_pfm_rows(const float* fb, int w, int h, float* out, hipStream_t st);

namespace {

using Clock = std::chrono::steady_clock;
using Bytes = std::vector<uint8_t>;

double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

int pool_threads(int requested) {
    if (requested > 0) return std::min(requested, 256);
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hw ? hw : 1u));
}

// Runs fn(0 .. n-1) on up to `threads` threads (the caller's among them); false if any task
// returned false.
bool parallel_for(int n, int threads, const std::function<bool(int)>& fn) {
    std::atomic<int> next{0};
    std::atomic<bool> ok{true};
    auto worker = [&] {
        for (int i = next++; i < n; i = next++)
            if (!fn(i)) ok = false;
    };
    const int k = std::max(1, std::min(threads, n));
    std::vector<std::thread> pool;
    pool.reserve(k - 1);
    for (int t = 1; t < k; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    return ok;
}

void put(Bytes& v, const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    v.insert(v.end(), b, b + n);
}
void put_str(Bytes& v, const char* s) { put(v, s, std::strlen(s) + 1); }  // with its terminating zero
template <class T>
void put_le(Bytes& v, T x) {  // the host is little-endian (x86-64)
    put(v, &x, sizeof(T));
}
void put_u32be(Bytes& v, uint32_t x) {
    const uint8_t b[4] = {(uint8_t)(x >> 24), (uint8_t)(x >> 16), (uint8_t)(x >> 8), (uint8_t)x};
    put(v, b, 4);
}

int zlevel(int level) { return level < 0 ? Z_DEFAULT_COMPRESSION : std::min(level, 9); }

// One zlib stream (RFC 1950) of src[0, n), deflated in pieces on the pool.
bool zlib_stream(const uint8_t* src, size_t n, int level, int threads, Bytes& out) {
    const size_t piece = std::max<size_t>(128 << 10, (n + threads - 1) / std::max(threads, 1));
    const int k = (int)std::max<size_t>(1, (n + piece - 1) / piece);
    std::vector<Bytes> part(k);
    std::vector<uLong> adler(k);
    const int lv = zlevel(level);
    bool ok = parallel_for(k, threads, [&](int i) {
        const size_t a = (size_t)i * piece, len = std::min(piece, n - a);
        z_stream zs;
        std::memset(&zs, 0, sizeof(zs));
        if (deflateInit2(&zs, lv, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) return false;
        if (a > 0) {
            const size_t d = std::min<size_t>(a, 32768);
            if (deflateSetDictionary(&zs, src + a - d, (uInt)d) != Z_OK) {
                deflateEnd(&zs);
                return false;
            }
        }
        part[i].resize(deflateBound(&zs, (uLong)len) + 64);
        zs.next_in = const_cast<Bytef*>(src + a);
        zs.avail_in = (uInt)len;
        zs.next_out = part[i].data();
        zs.avail_out = (uInt)part[i].size();
        const bool last = i == k - 1;
        const int rc = deflate(&zs, last ? Z_FINISH : Z_SYNC_FLUSH);
        const bool good = last ? rc == Z_STREAM_END : (rc == Z_OK && zs.avail_in == 0 && zs.avail_out > 0);
        part[i].resize(zs.total_out);
        deflateEnd(&zs);
        adler[i] = adler32(adler32(0L, Z_NULL, 0), src + a, (uInt)len);
        return good;
    });
    if (!ok) return false;
    // CMF 0x78 (deflate, 32 KiB window); FLG's level hint, with the check bits the format wants
    const uint8_t flg = lv == 0 || lv == 1 ? 0x01 : (lv >= 2 && lv <= 5 ? 0x5e : (lv >= 7 ? 0xda : 0x9c));
    out.push_back(0x78);
    out.push_back(flg);
    uLong ad = adler32(0L, Z_NULL, 0);
    for (int i = 0; i < k; ++i) {
        put(out, part[i].data(), part[i].size());
        const size_t len = std::min(piece, n - (size_t)i * piece);
        ad = adler32_combine(ad, adler[i], (z_off_t)len);
    }
    put_u32be(out, (uint32_t)ad);
    return true;
}

void png_chunk(Bytes& f, const char* type, const uint8_t* data, size_t n) {
    put_u32be(f, (uint32_t)n);
    const size_t at = f.size();
    put(f, type, 4);
    if (n) put(f, data, n);
    put_u32be(f, (uint32_t)crc32(0L, f.data() + at, (uInt)(n + 4)));
}

// PNG signature, IHDR (8-bit RGBA), one IDAT with the filtered rows' zlib stream, IEND.
bool png_file(const uint8_t* rows, int w, int h, int level, int threads, Bytes& f) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    put(f, sig, 8);
    Bytes ihdr;
    put_u32be(ihdr, (uint32_t)w);
    put_u32be(ihdr, (uint32_t)h);
    const uint8_t tail[5] = {8, 6, 0, 0, 0};  // bit depth, colour type RGBA, deflate, filters, no interlace
    put(ihdr, tail, 5);
    png_chunk(f, "IHDR", ihdr.data(), ihdr.size());
    Bytes z;
    if (!zlib_stream(rows, (size_t)h * (1 + 4 * (size_t)w), level, threads, z)) return false;
    // a chunk holds at most 2^31 - 1 bytes
    for (size_t a = 0; a < z.size(); a += 0x40000000u)
        png_chunk(f, "IDAT", z.data() + a, std::min<size_t>(0x40000000u, z.size() - a));
    png_chunk(f, "IEND", nullptr, 0);
    return true;
}

void exr_attr(Bytes& f, const char* name, const char* type, const Bytes& value) {
    put_str(f, name);
    put_str(f, type);
    put_le<int32_t>(f, (int32_t)value.size());
    put(f, value.data(), value.size());
}

// OpenEXR single-part scanline file: magic, version 2, the header attributes, the chunk offset
// table, then per 16-line block (y, size, data). pre holds the blocks after the ZIP pre-pass at
// a stride of 128 w bytes; a block that deflate does not shrink is stored as its plain halves,
// recovered here by undoing the predictor and the byte split.
bool exr_file(const uint8_t* pre, int w, int h, int level, int threads, Bytes& f) {
    const uint8_t magic[8] = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};
    put(f, magic, 8);
    Bytes v;
    for (const char* ch : {"A", "B", "G", "R"}) {
        put_str(v, ch);
        put_le<int32_t>(v, 1);  // HALF
        const uint8_t linear_reserved[4] = {0, 0, 0, 0};
        put(v, linear_reserved, 4);
        put_le<int32_t>(v, 1);  // x sampling
        put_le<int32_t>(v, 1);  // y sampling
    }
    v.push_back(0);
    exr_attr(f, "channels", "chlist", v);
    exr_attr(f, "compression", "compression", Bytes{3});  // ZIP, 16 scanlines
    v.clear();
    for (int32_t x : {0, 0, w - 1, h - 1}) put_le<int32_t>(v, x);
    exr_attr(f, "dataWindow", "box2i", v);
    exr_attr(f, "displayWindow", "box2i", v);
    exr_attr(f, "lineOrder", "lineOrder", Bytes{0});  // increasing y
    v.clear();
    put_le<float>(v, 1.0f);
    exr_attr(f, "pixelAspectRatio", "float", v);
    v.clear();
    put_le<float>(v, 0.0f);
    put_le<float>(v, 0.0f);
    exr_attr(f, "screenWindowCenter", "v2f", v);
    v.clear();
    put_le<float>(v, 1.0f);
    exr_attr(f, "screenWindowWidth", "float", v);
    f.push_back(0);  // end of header

    const int nb = (h + 15) / 16;
    const size_t stride = 128 * (size_t)w;
    std::vector<Bytes> blk(nb);
    const int lv = zlevel(level);
    bool ok = parallel_for(nb, threads, [&](int b) {
        const int lines = std::min(16, h - 16 * b);
        const size_t S = 8 * (size_t)w * lines;
        const uint8_t* t = pre + (size_t)b * stride;
        uLongf cap = compressBound((uLong)S);
        blk[b].resize(cap);
        if (compress2(blk[b].data(), &cap, t, (uLong)S, lv) != Z_OK) return false;
        if (cap < S) {
            blk[b].resize(cap);
            return true;
        }
        // stored plain: r[j] = t[j] + r[j-1] - 128 (mod 256), then low bytes r[0, m) and high
        // bytes r[m, 2m) interleaved back into halves
        Bytes r(S);
        r[0] = t[0];
        for (size_t j = 1; j < S; ++j) r[j] = (uint8_t)(t[j] + r[j - 1] - 128);
        blk[b].resize(S);
        const size_t m = S / 2;
        for (size_t i = 0; i < m; ++i) {
            blk[b][2 * i] = r[i];
            blk[b][2 * i + 1] = r[m + i];
        }
        return true;
    });
    if (!ok) return false;
    const size_t table = f.size();
    f.resize(table + 8 * (size_t)nb);
    for (int b = 0; b < nb; ++b) {
        const uint64_t at = f.size();
        std::memcpy(f.data() + table + 8 * (size_t)b, &at, 8);
        put_le<int32_t>(f, 16 * b);
        put_le<int32_t>(f, (int32_t)blk[b].size());
        put(f, blk[b].data(), blk[b].size());
    }
    return true;
}

size_t payload_bytes(int fmt, int w, int h) {
    const size_t n = (size_t)w * h;
    switch (fmt) {
        case SIRIUS_IMAGE_PPM:
        case SIRIUS_IMAGE_PPM_SRGB: return 3 * n;
        case SIRIUS_IMAGE_PNG: return (size_t)h * (1 + 4 * (size_t)w);
        case SIRIUS_IMAGE_EXR: return 8 * n;
        case SIRIUS_IMAGE_PFM: return 12 * n;
        default: return 16 * n;
    }
}

int encode(Handle& H, const float* rgba, int w, int h, int on_device, const SiriusHipImageOpts& o,
           SiriusHipImageStats* st) {
    const auto t0 = Clock::now();
    const int fmt = o.format;
    if (fmt < SIRIUS_IMAGE_PPM || fmt > SIRIUS_IMAGE_PPM_SRGB) return H.fail(SIRIUS_HIP_E_INVALID, "image format");
    if (!(o.exposure == o.exposure) || o.zlib_level > 9 || o.zlib_level < -1)
        return H.fail(SIRIUS_HIP_E_INVALID, "image options (exposure, zlib_level in -1..9)");
    int rc = H.ensure_device();
    if (rc) return rc;
    const float* src = rgba;
    if (!rgba) {
        if (!H.initialised || !H.fb_dev) return H.fail(SIRIUS_HIP_E_STATE, "no framebuffer (initialise and launch first)");
        if ((w || h) && (w != H.fb_w || h != H.fb_h))
            return H.fail(SIRIUS_HIP_E_INVALID, "width and height must be 0 or the framebuffer's");
        w = H.fb_w;
        h = H.fb_h;
        src = H.fb_dev;
    } else if (w <= 0 || h <= 0) {
        return H.fail(SIRIUS_HIP_E_INVALID, "empty image");
    }
    if ((int64_t)w * h > (int64_t)1 << 30) return H.fail(SIRIUS_HIP_E_INVALID, "image larger than 2^30 pixels");
    const size_t n = (size_t)w * h;
    hipError_t e = hipSuccess;
    if (rgba && !on_device) {
        float* d = (float*)H.scratch("img_src", 16 * n, &rc);
        if (rc) return rc;
        if ((e = hipMemcpyAsync(d, rgba, 16 * n, hipMemcpyHostToDevice, H.stream)) != hipSuccess)
            return H.hip_fail(e, "image upload");
        src = d;
    }
    const size_t pay = payload_bytes(fmt, w, h);
    if (H.pinned_bytes < pay) {
        if (H.pinned) (void)hipHostFree(H.pinned);
        H.pinned = nullptr;
        H.pinned_bytes = 0;
        if ((e = hipHostMalloc((void**)&H.pinned, pay, hipHostMallocDefault)) != hipSuccess) {
            H.pinned = nullptr;
            return H.hip_fail(e, "pinned staging");
        }
        H.pinned_bytes = pay;
    }
    (void)hipEventRecord(H.ev0, H.stream);
    const void* dev_pay = src;  // raw: the frame itself
    if (fmt == SIRIUS_IMAGE_PPM || fmt == SIRIUS_IMAGE_PPM_SRGB) {
        uint8_t* d = (uint8_t*)H.scratch("img_pay", pay, &rc);
        if (rc) return rc;
        e = launch_ldr_pixels(fmt == SIRIUS_IMAGE_PPM ? 0 : 1, src, (int64_t)n, o.exposure, d, H.stream);
        dev_pay = d;
    } else if (fmt == SIRIUS_IMAGE_PNG) {
        uint8_t* px = (uint8_t*)H.scratch("img_px", 4 * n, &rc);
        if (rc) return rc;
        uint8_t* d = (uint8_t*)H.scratch("img_pay", pay, &rc);
        if (rc) return rc;
        e = launch_ldr_pixels(2, src, (int64_t)n, 1.0f, px, H.stream);
        if (e == hipSuccess) e = launch_png_filter(px, w, h, 4, d, H.stream);
        dev_pay = d;
    } else if (fmt == SIRIUS_IMAGE_EXR) {
        // blocks at a stride of 128 w bytes: the last, shorter block ends inside the payload
        uint8_t* d = (uint8_t*)H.scratch("img_pay", pay, &rc);
        if (rc) return rc;
        e = launch_exr_zip_prepass(src, w, h, d, H.stream);
        dev_pay = d;
    } else if (fmt == SIRIUS_IMAGE_PFM) {
        float* d = (float*)H.scratch("img_pay", pay, &rc);
        if (rc) return rc;
        e = launch_pfm_rows(src, w, h, d, H.stream);
        dev_pay = d;
    }
    if (e != hipSuccess) return H.hip_fail(e, "image conversion");
    (void)hipEventRecord(H.ev1, H.stream);
    if ((e = hipMemcpyAsync(H.pinned, dev_pay, pay, hipMemcpyDeviceToHost, H.stream)) != hipSuccess)
        return H.hip_fail(e, "image copy");
    (void)hipEventRecord(H.ev2, H.stream);
    if ((e = hipEventSynchronize(H.ev2)) != hipSuccess) return H.hip_fail(e, "image synchronize");
    float conv = 0, copy = 0;
    (void)hipEventElapsedTime(&conv, H.ev0, H.ev1);
    (void)hipEventElapsedTime(&copy, H.ev1, H.ev2);

    const auto t1 = Clock::now();
    Bytes& f = H.image;
    f.clear();
    const int threads = pool_threads(o.threads);
    char head[64];
    bool ok = true;
    switch (fmt) {
        case SIRIUS_IMAGE_PPM:
        case SIRIUS_IMAGE_PPM_SRGB:
            std::snprintf(head, sizeof head, "P6\n%d %d\n255\n", w, h);
            f.reserve(std::strlen(head) + pay);
            put(f, head, std::strlen(head));
            put(f, H.pinned, pay);
            break;
        case SIRIUS_IMAGE_PFM:
            std::snprintf(head, sizeof head, "PF\n%d %d\n-1.0\n", w, h);
            f.reserve(std::strlen(head) + pay);
            put(f, head, std::strlen(head));
            put(f, H.pinned, pay);
            break;
        case SIRIUS_IMAGE_RAW:
            f.reserve(12 + pay);
            put_le<int32_t>(f, w);
            put_le<int32_t>(f, h);
            put_le<int32_t>(f, 4);
            put(f, H.pinned, pay);
            break;
        case SIRIUS_IMAGE_PNG: ok = png_file(H.pinned, w, h, o.zlib_level, threads, f); break;
        default: ok = exr_file(H.pinned, w, h, o.zlib_level, threads, f); break;
    }
    if (!ok) return H.fail(SIRIUS_HIP_E_NOMEM, "compression failed");
    if (st) {
        st->convert_ms = conv;
        st->copy_ms = copy;
        st->pack_ms = ms_since(t1);
        st->total_ms = ms_since(t0);
        st->payload_bytes = pay;
        st->file_bytes = f.size();
    }
    return SIRIUS_HIP_OK;
}

bool write_file(const char* path, const void* data, size_t n) {
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(data, 1, n, fp) == n;
    return std::fclose(fp) == 0 && ok;
}

}  // namespace

}  // namespace sirius

using namespace sirius;

#define GUARD_BEGIN try {
#define GUARD_END(h)                                                                     \
    }                                                                                    \
    catch (const std::bad_alloc&) { return (h)->fail(SIRIUS_HIP_E_NOMEM, "host allocation"); } \
    catch (...) { return (h)->fail(SIRIUS_HIP_E_DEVICE, "unexpected exception"); }

extern "C" {

void sirius_hip_default_image_opts(SiriusHipImageOpts* o, int format) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->format = format;
    o->exposure = 1.0f;
    o->zlib_level = -1;
    o->threads = 0;
}

int sirius_hip_encode_image(SiriusHipHandle hh, const float* rgba, int width, int height, int on_device,
                            const SiriusHipImageOpts* o, const unsigned char** data, size_t* size,
                            SiriusHipImageStats* stats) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
        if (!o || !data || !size) return H->fail(SIRIUS_HIP_E_INVALID, "opts, data and size are required");
        *data = nullptr;
        *size = 0;
        int rc = encode(*H, rgba, width, height, on_device, *o, stats);
        if (rc) return rc;
        *data = H->image.data();
        *size = H->image.size();
        return SIRIUS_HIP_OK;
    GUARD_END(H)
}

int sirius_hip_write_image(SiriusHipHandle hh, const char* path, const float* rgba, int width, int height,
                           int on_device, const SiriusHipImageOpts* o, SiriusHipImageStats* stats) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
        if (!path || !o) return H->fail(SIRIUS_HIP_E_INVALID, "path and opts are required");
        const auto t0 = Clock::now();
        int rc = encode(*H, rgba, width, height, on_device, *o, stats);
        if (rc) return rc;
        if (!write_file(path, H->image.data(), H->image.size()))
            return H->fail(SIRIUS_HIP_E_IO, std::string("cannot write ") + path);
        if (stats) stats->total_ms = ms_since(t0);
        return SIRIUS_HIP_OK;
    GUARD_END(H)
}

int sirius_hip_write_rays(SiriusHipHandle hh, const char* path, uint32_t fields, const SiriusHipCamera* cams,
                          int n_cams, const SiriusHipMetric* metrics, int n_metrics, const SiriusHipTraceOpts* opts) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
        if (!path) return H->fail(SIRIUS_HIP_E_INVALID, "path is required");
        if (fields & ~SIRIUS_RAYS_ALL) return H->fail(SIRIUS_HIP_E_INVALID, "unknown SIRIUS_RAYS_* bits");
        if (n_cams < 0 || n_metrics < 0 || (n_cams && !cams) || (n_metrics && !metrics))
            return H->fail(SIRIUS_HIP_E_INVALID, "camera and metric records");
        if (!H->last) return H->fail(SIRIUS_HIP_E_STATE, "no trace results to write");
        const DevOut& o = *H->last;
        const uint64_t n = (uint64_t)o.n_rays;
        const int nb = std::max(H->last_bins, 1);
        struct Field { uint32_t bit; const void* src; size_t bytes; } fl[] = {
            {SIRIUS_RAYS_CLS, o.cls, n},
            {SIRIUS_RAYS_FLAGS, o.flags, n},
            {SIRIUS_RAYS_THETA, o.theta, 4 * n},
            {SIRIUS_RAYS_PHI, o.phi, 4 * n},
            {SIRIUS_RAYS_LAMBDA, o.lam, 4 * n},
            {SIRIUS_RAYS_DISK_R, o.disk_r, 4 * n},
            {SIRIUS_RAYS_REDSHIFT, o.redshift, 4 * n},
            {SIRIUS_RAYS_NULL_DRIFT, o.null_drift, 4 * n},
            {SIRIUS_RAYS_STEPS, o.steps, 4 * n},
            {SIRIUS_RAYS_INTENSITY, o.intensity, 4 * n * (size_t)nb},
            {SIRIUS_RAYS_STATE, o.state, 64 * n},
        };
        uint32_t present = 0;
        for (auto& x : fl)
            if ((fields & x.bit) && x.src) present |= x.bit;
        Bytes head(128, 0);
        std::memcpy(head.data(), "SIRRAYS1", 8);
        auto at = [&](size_t off, auto v) { std::memcpy(head.data() + off, &v, sizeof(v)); };
        at(8, (uint32_t)128);
        at(12, (uint32_t)nb);
        at(16, n);
        at(24, present);
        at(28, (int32_t)n_cams);
        at(32, (uint32_t)sizeof(SiriusHipCamera));
        at(36, (int32_t)n_metrics);
        at(40, (uint32_t)sizeof(SiriusHipMetric));
        at(44, (uint32_t)(opts ? sizeof(SiriusHipTraceOpts) : 0));
        FILE* fp = std::fopen(path, "wb");
        if (!fp) return H->fail(SIRIUS_HIP_E_IO, std::string("cannot write ") + path);
        bool ok = std::fwrite(head.data(), 1, head.size(), fp) == head.size();
        size_t pos = head.size();
        auto emit = [&](const void* p, size_t bytes) {
            if (ok && bytes) ok = std::fwrite(p, 1, bytes, fp) == bytes;
            pos += bytes;
        };
        emit(cams, sizeof(SiriusHipCamera) * (size_t)n_cams);
        emit(metrics, sizeof(SiriusHipMetric) * (size_t)n_metrics);
        if (opts) emit(opts, sizeof(SiriusHipTraceOpts));
        const uint8_t zeros[8] = {0};
        size_t biggest = 0;
        for (auto& x : fl)
            if (present & x.bit) biggest = std::max(biggest, x.bytes);
        std::vector<uint8_t> stage(biggest);
        for (auto& x : fl) {
            if (!(present & x.bit)) continue;
            emit(zeros, (8 - pos % 8) % 8);
            hipError_t e = hipMemcpyAsync(stage.data(), x.src, x.bytes, hipMemcpyDeviceToHost, H->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(H->stream);
            if (e != hipSuccess) {
                std::fclose(fp);
                return H->hip_fail(e, "ray results copy");
            }
            emit(stage.data(), x.bytes);
        }
        ok = (std::fclose(fp) == 0) && ok;
        if (!ok) return H->fail(SIRIUS_HIP_E_IO, std::string("cannot write ") + path);
        return SIRIUS_HIP_OK;
    GUARD_END(H)
}

}  // extern "C"
