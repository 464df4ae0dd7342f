// image_kernels.hip — device side of the image writers: the per-pixel conversions of a float
// RGBA framebuffer (the frame stays in HBM; only the encoded payload crosses PCIe).
//
// Conversions, each following the reference writer whose files they must decode to:
//   sRGB 8-bit          PNGWriter::toSRGB8 (Sirius.Render/Output/OUPN001A.h:103-112)
//   ACES + 1/2.2 gamma  BufferWriter::writePPM (Sirius.Render/Buffer/BFIO001A.h:29-36,44-68)
//   float -> half       tinyexr's float_to_half_full (lib/tinyexr/tinyexr.h:985-1023), which
//                       EXRWriter (OUEW001A.cpp:109-178) asks for: magnitude rounded half-up,
//                       float denormals to zero, overflow to infinity
// Every float step is a separate IEEE operation in the reference's order (contraction is off in
// this file's conversions and the division is the IEEE one: this file is compiled without the
// approximate FP32 division); pow is evaluated in FP64 and rounded once to float.
//
// Container work that is data-parallel also runs here: the PNG row filters (per row, the filter
// with the least sum of |signed bytes|, the usual heuristic) and the ZIP pre-pass of each
// 16-line EXR block (the OpenEXR byte split into low and high halves and the delta predictor),
// computed directly from the half values, one thread per half.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sirius {

namespace {

#pragma clang fp contract(off)

__device__ __forceinline__ float powf_rn(float x, float y) { return (float)pow((double)x, (double)y); }

// float -> uint8 as a C++ cast of a value in [0, 255] (truncation); NaN, whose cast the C++
// reference leaves undefined (x86 gives 0), is 0
__device__ __forceinline__ uint8_t to_u8(float v) { return v == v ? (uint8_t)(uint32_t)v : (uint8_t)0; }

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

__device__ __forceinline__ uint8_t srgb8(float linear) {
    const float c = clampf(linear, 0.0f, 1.0f);
    const float s = c <= 0.0031308f ? 12.92f * c : 1.055f * powf_rn(c, 1.0f / 2.4f) - 0.055f;
    return to_u8(clampf(s * 255.0f + 0.5f, 0.0f, 255.0f));
}

__device__ __forceinline__ uint8_t alpha8(float a) { return to_u8(clampf(a * 255.0f + 0.5f, 0.0f, 255.0f)); }

__device__ __forceinline__ float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return clampf((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}

__device__ __forceinline__ uint8_t aces8(float x, float exposure) {
    const float t = powf_rn(aces(x * exposure), 1.0f / 2.2f);
    return to_u8(clampf(t * 255.0f, 0.0f, 255.0f));
}

// Half bits of a float: sign kept; exponent 0 (zero, denormal) -> 0; exponent 255 -> infinity,
// or the quiet NaN 0x7e00 with the sign; exponents past the half range -> infinity; below it
// the significand (with its hidden bit) is shifted into a half denormal; otherwise the top 10
// significand bits. Where bits are dropped, the result is incremented when the first dropped
// bit is set (a carry may run into the exponent, up to infinity).
__device__ __forceinline__ uint32_t half_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const int exp = (int)((u >> 23) & 0xffu);
    const uint32_t mant = u & 0x7fffffu;
    uint32_t h = 0;
    if (exp == 0) {
        h = 0;
    } else if (exp == 255) {
        h = 0x7c00u | (mant ? 0x200u : 0u);
    } else {
        const int e = exp - 127 + 15;
        if (e >= 31) {
            h = 0x7c00u;
        } else if (e <= 0) {
            const int shift = 14 - e;  // 14 .. ; no bit survives past 24
            if (shift <= 24) {
                const uint32_t m = mant | 0x800000u;
                h = m >> shift;
                if ((m >> (shift - 1)) & 1u) ++h;
            }
        } else {
            h = ((uint32_t)e << 10) | (mant >> 13);
            if (mant & 0x1000u) ++h;
        }
    }
    return (sign | h) & 0xffffu;
}

// EXR channel slot c (the file's alphabetical A, B, G, R) -> framebuffer component (R G B A)
__device__ __forceinline__ float exr_channel(const float4& p, int slot) {
    return slot == 0 ? p.w : (slot == 1 ? p.z : (slot == 2 ? p.y : p.x));
}

}  // namespace

// MODE 0: ACES PPM, 3 bytes per pixel; 1: sRGB PPM, 3 bytes; 2: sRGB + alpha (PNG), 4 bytes.
template <int MODE>
__global__ __launch_bounds__(256) void k_ldr_pixels(const float4* __restrict__ fb, int64_t n, float exposure,
                                                    uint8_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = fb[i];
    if (MODE == 0) {
        out[3 * i + 0] = aces8(p.x, exposure);
        out[3 * i + 1] = aces8(p.y, exposure);
        out[3 * i + 2] = aces8(p.z, exposure);
    } else if (MODE == 1) {
        out[3 * i + 0] = srgb8(p.x);
        out[3 * i + 1] = srgb8(p.y);
        out[3 * i + 2] = srgb8(p.z);
    } else {
        uchar4 q;
        q.x = srgb8(p.x);
        q.y = srgb8(p.y);
        q.z = srgb8(p.z);
        q.w = alpha8(p.w);
        reinterpret_cast<uchar4*>(out)[i] = q;
    }
}

namespace {

__device__ __forceinline__ int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

__device__ __forceinline__ uint8_t png_filter(int f, int x, int a, int b, int c) {
    switch (f) {
        case 0: return (uint8_t)x;
        case 1: return (uint8_t)(x - a);
        case 2: return (uint8_t)(x - b);
        case 3: return (uint8_t)(x - ((a + b) >> 1));
        default: return (uint8_t)(x - paeth(a, b, c));
    }
}

}  // namespace

// One workgroup per row: the five filtered forms' costs (sum of |signed byte|), the cheapest
// (the lowest filter number on a tie), then the row written as its filter byte and bytes.
// px: rows of w*bpp bytes; out: rows of 1 + w*bpp bytes.
__global__ __launch_bounds__(256) void k_png_filter(const uint8_t* __restrict__ px, int w, int h, int bpp,
                                                   uint8_t* __restrict__ out) {
    __shared__ uint32_t cost[5][256];
    const int row = blockIdx.x;
    if (row >= h) return;
    const int64_t len = (int64_t)w * bpp;
    const uint8_t* cur = px + (int64_t)row * len;
    const uint8_t* up = row > 0 ? cur - len : nullptr;
    uint32_t s[5] = {0, 0, 0, 0, 0};
    for (int64_t k = threadIdx.x; k < len; k += blockDim.x) {
        const int x = cur[k];
        const int a = k >= bpp ? cur[k - bpp] : 0;
        const int b = up ? up[k] : 0;
        const int c = (up && k >= bpp) ? up[k - bpp] : 0;
        for (int f = 0; f < 5; ++f) s[f] += (uint32_t)abs((int)(int8_t)png_filter(f, x, a, b, c));
    }
    for (int f = 0; f < 5; ++f) cost[f][threadIdx.x] = s[f];
    __syncthreads();
    for (int stride = blockDim.x / 2; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride)
            for (int f = 0; f < 5; ++f) cost[f][threadIdx.x] += cost[f][threadIdx.x + stride];
        __syncthreads();
    }
    int best = 0;
    for (int f = 1; f < 5; ++f)
        if (cost[f][0] < cost[best][0]) best = f;
    uint8_t* o = out + (int64_t)row * (len + 1);
    if (threadIdx.x == 0) o[0] = (uint8_t)best;
    for (int64_t k = threadIdx.x; k < len; k += blockDim.x) {
        const int x = cur[k];
        const int a = k >= bpp ? cur[k - bpp] : 0;
        const int b = up ? up[k] : 0;
        const int c = (up && k >= bpp) ? up[k - bpp] : 0;
        o[1 + k] = png_filter(best, x, a, b, c);
    }
}

// EXR ZIP pre-pass. Block b holds lines 16b .. 16b+L-1 (L = 16, or what is left), each line the
// channels A, B, G, R of w halves (little-endian): m = 4 w L halves, 2m bytes. The OpenEXR
// reorder puts every even byte (the halves' low bytes) first and every odd byte (high bytes)
// after them; the predictor replaces each byte after the first by (byte - previous + 128) mod
// 256. Thread i of a block writes both bytes that come from half i. Blocks are laid out at a
// stride of 128 w bytes.
__global__ __launch_bounds__(256) void k_exr_zip_prepass(const float4* __restrict__ fb, int w, int h,
                                                        uint8_t* __restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per_line = 4 * (int64_t)w;
    const int64_t total = per_line * h;
    if (g >= total) return;
    const int64_t per_block = 16 * per_line;
    const int64_t blk = g / per_block;
    const int64_t i = g - blk * per_block;
    const int lines = min(16, h - (int)(blk * 16));
    const int64_t m = per_line * lines;
    auto half_at = [&](int64_t j) -> uint32_t {
        const int64_t line = j / per_line;
        const int64_t rem = j - line * per_line;
        const int slot = (int)(rem / w);
        const int64_t x = rem - (int64_t)slot * w;
        const float4 p = fb[(blk * 16 + line) * (int64_t)w + x];
        return half_bits(exr_channel(p, slot));
    };
    const uint32_t hi = half_at(i);
    const uint32_t prev = i > 0 ? half_at(i - 1) : half_at(m - 1);
    uint8_t* o = out + blk * per_block * 2;
    o[i] = (uint8_t)(i == 0 ? (hi & 0xffu) : ((hi & 0xffu) - (prev & 0xffu) + 128u));
    // the high-byte run starts after the last low byte (of half m-1)
    o[m + i] = (uint8_t)((hi >> 8) - (i == 0 ? (prev & 0xffu) : (prev >> 8)) + 128u);
}

// PFM payload: RGB floats, rows bottom to top (EXRWriter::writePFM, OUEW001A.h:110-126).
__global__ __launch_bounds__(256) void k_pfm_rows(const float4* __restrict__ fb, int w, int h, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int64_t y = i / w, x = i - y * w;
    const float4 p = fb[(h - 1 - y) * (int64_t)w + x];
    out[3 * i + 0] = p.x;
    out[3 * i + 1] = p.y;
    out[3 * i + 2] = p.z;
}

static unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_ldr_pixels(int mode, const float* fb, int64_t n, float exposure, uint8_t* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const float4* f = reinterpret_cast<const float4*>(fb);
    if (mode == 0)
        hipLaunchKernelGGL(k_ldr_pixels<0>, dim3(grid_of(n)), dim3(256), 0, st, f, n, exposure, out);
    else if (mode == 1)
        hipLaunchKernelGGL(k_ldr_pixels<1>, dim3(grid_of(n)), dim3(256), 0, st, f, n, exposure, out);
    else
        hipLaunchKernelGGL(k_ldr_pixels<2>, dim3(grid_of(n)), dim3(256), 0, st, f, n, exposure, out);
    return hipGetLastError();
}

hipError_t launch_png_filter(const uint8_t* px, int w, int h, int bpp, uint8_t* out, hipStream_t st) {
    if (w <= 0 || h <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_png_filter, dim3((unsigned)h), dim3(256), 0, st, px, w, h, bpp, out);
    return hipGetLastError();
}

hipError_t launch_exr_zip_prepass(const float* fb, int w, int h, uint8_t* out, hipStream_t st) {
    const int64_t n = 4 * (int64_t)w * h;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_exr_zip_prepass, dim3(grid_of(n)), dim3(256), 0, st, reinterpret_cast<const float4*>(fb), w,
                       h, out);
    return hipGetLastError();
}

hipError_t launch_pfm_rows(const float* fb, int w, int h, float* out, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pfm_rows, dim3(grid_of(n)), dim3(256), 0, st, reinterpret_cast<const float4*>(fb), w, h, out);
    return hipGetLastError();
}

}  // namespace sirius
