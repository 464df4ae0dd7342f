// util_kernels.hip — small kernels around the tracer: statistics reduction, distance of each
// ray from the critical curve (the surrogate's fallback band), and RGBA shading of the
// framebuffer that IAccelerator::getFrameBuffer returns.
#include <hip/hip_runtime.h>

#include "../../include/sirius_hip.h"
#include "geodesic.hpp"
#include "sirius_bundle.hpp"

namespace sirius {

// acc[0] attempts, acc[1] rejected, acc[2] landings, acc[3] max null drift (double bits),
// acc[4] escalated rays, acc[5] surrogate rays, acc[6] audited rays, acc[7] attempts of the
// escalated rays. Grid-stride loop, wave then block reduction in LDS, one atomic per block
// and counter.
// hist (SIRIUS_DRIFT_BINS counters): the null drift of every ray binned by log10, 1/16 decade per
// bin from 1e-20 (bin 0 also holds smaller values and zero) to 1 (the last bin holds larger
// values), for its median; per-block counts in LDS, one atomic per nonzero bin per block.
__global__ __launch_bounds__(256) void k_reduce_stats(DevOut o, int64_t n, unsigned long long* acc,
                                                      uint32_t* hist) {
    __shared__ uint32_t h[SIRIUS_DRIFT_BINS];
    for (int b = threadIdx.x; b < SIRIUS_DRIFT_BINS; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double dmax = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t st = o.steps[i];
        v[0] += st;
        v[1] += o.rejected[i];
        v[2] += o.landings[i];
        uint8_t f = o.flags[i];
        v[4] += (f & SIRIUS_FLAG_ESCALATED) ? 1 : 0;
        v[7] += (f & SIRIUS_FLAG_ESCALATED) ? st : 0;
        v[5] += (f & SIRIUS_FLAG_SURROGATE) ? 1 : 0;
        v[6] += (f & SIRIUS_FLAG_AUDITED) ? 1 : 0;
        double d = (double)o.null_drift[i];
        dmax = d > dmax ? d : dmax;
        float lb = d > 0.0 ? (log10f((float)d) + 20.0f) * 16.0f : 0.0f;
        int b = lb > 0.0f ? (int)lb : 0;
        atomicAdd(&h[b < SIRIUS_DRIFT_BINS ? b : SIRIUS_DRIFT_BINS - 1], 1u);
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 8; ++k) v[k] += __shfl_xor(v[k], off);
        double t = __shfl_xor(dmax, off);
        dmax = t > dmax ? t : dmax;
    }
    __shared__ unsigned long long part[4][8];
    __shared__ double pmax[4];
    int w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        for (int k = 0; k < 8; ++k) part[w][k] = v[k];
        pmax[w] = dmax;
    }
    __syncthreads();
    if (threadIdx.x < 8 && threadIdx.x != 3) {
        unsigned long long t = 0;
        for (int j = 0; j < (int)(blockDim.x >> 6); ++j) t += part[j][threadIdx.x];
        atomicAdd(&acc[threadIdx.x], t);
    } else if (threadIdx.x == 3) {
        double t = 0.0;
        for (int j = 0; j < (int)(blockDim.x >> 6); ++j) t = pmax[j] > t ? pmax[j] : t;
        atomicMax(&acc[3], (unsigned long long)__double_as_longlong(t));
    }
    for (int b = threadIdx.x; b < SIRIUS_DRIFT_BINS; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

hipError_t launch_reduce_stats(const DevOut& o, int64_t n, unsigned long long* acc, uint32_t* hist,
                               hipStream_t st) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 512) blocks = 512;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_reduce_stats, dim3((unsigned)blocks), dim3(256), 0, st, o, n, acc, hist);
    return hipGetLastError();
}

// Critical-curve distance of every ray (radial, from the per-camera tables; 1e30 for metrics
// without a photon sphere).
__global__ __launch_bounds__(256) void k_critical_distance(const CamConst* __restrict__ cams, int n_cam,
                                                           const float* __restrict__ tabs, int64_t n,
                                                           float* __restrict__ out) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    int lo = 0, hi = n_cam - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cams[mid].offset <= id) lo = mid; else hi = mid - 1;
    }
    const CamConst& c = cams[lo];
    if (c.family != 0) { out[id] = 1e30f; return; }
    int64_t pix = id - c.offset;
    Spacetime<double> s = make_spacetime<double>(c);
    double x[4], p[4];
    init_state<double>(c, s, (int)(pix % c.width), (int)(pix / c.width), x, p);
    out[id] = (float)crit_distance<double>(tabs + (int64_t)lo * (SIRIUS_CC_TABLE + 1), s, x, p);
}

hipError_t launch_critical_distance(const CamConst* cams, int n_cam, const float* tabs, int64_t n, float* out,
                                    hipStream_t st) {
    hipLaunchKernelGGL(k_critical_distance, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cams, n_cam, tabs,
                       n, out);
    return hipGetLastError();
}

// Gate of the FP32 path: rays within `band` of the critical curve go to list B (traced in FP64
// from the start, on a second stream), the others within `long_band` to list L (the FP32
// tracer takes them first: they wind near the photon region and are the longest) and all
// others to list A; rays within `crit_band` get SIRIUS_FLAG_CRITICAL in flags[] (every ray's
// flag byte is written). counts[0..2] = |A|, |B|, |L|. A block
// of 256 threads classifies a contiguous chunk of GATE_RAYS rays and reserves its slots with
// one atomic per list (device-scope atomics on one address serialise, so one per wave is too
// many at 1080p); within a chunk the lists keep ray order.
constexpr int GATE_PASSES = 8;
constexpr int GATE_RAYS = 256 * GATE_PASSES;

// |distance| of ray id from its camera's critical curve (M); +inf for other spacetimes
__device__ __forceinline__ float gate_distance(const CamConst* __restrict__ cams, int n_cam,
                                              const float* __restrict__ tabs, int64_t id) {
    int lo = 0, hi = n_cam - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cams[mid].offset <= id) lo = mid; else hi = mid - 1;
    }
    const CamConst& c = cams[lo];
    if (c.family != 0) return INFINITY;
    // FP32 is ample for band membership (alpha, beta ~ 10 M, band ~ 0.1 M)
    int64_t pix = id - c.offset;
    Spacetime<float> s = make_spacetime<float>(c);
    float x[4], p[4];
    init_state<float>(c, s, (int)(pix % c.width), (int)(pix / c.width), x, p);
    return fabsf(crit_distance<float>(tabs + (int64_t)lo * (SIRIUS_CC_TABLE + 1), s, x, p));
}

__global__ __launch_bounds__(256) void k_gate(const CamConst* __restrict__ cams, int n_cam,
                                              const float* __restrict__ tabs, int64_t n, float band,
                                              float crit_band, float long_band, int32_t* __restrict__ list_a,
                                              int32_t* __restrict__ list_b, int32_t* __restrict__ list_l,
                                              uint32_t* __restrict__ counts, uint8_t* __restrict__ flags) {
    __shared__ uint32_t off_a[GATE_PASSES][4], off_b[GATE_PASSES][4], off_l[GATE_PASSES][4];
    __shared__ uint32_t base[3];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const int64_t chunk = (int64_t)blockIdx.x * GATE_RAYS;
    const unsigned long long below = lane ? ((1ull << lane) - 1ull) : 0ull;
    uint32_t kinds = 0;  // 2 bits per pass: 0 list A, 1 list B (band), 2 list L (long), 3 invalid
#pragma unroll
    for (int k = 0; k < GATE_PASSES; ++k) {
        int64_t id = chunk + k * 256 + threadIdx.x;
        bool valid = id < n;
        float d = valid ? gate_distance(cams, n_cam, tabs, id) : INFINITY;
        uint32_t kind = !valid ? 3u : (d < band ? 1u : (d < long_band ? 2u : 0u));
        if (valid) flags[id] = d < crit_band ? (uint8_t)SIRIUS_FLAG_CRITICAL : (uint8_t)0;
        kinds |= kind << (2 * k);
        unsigned long long ma = __ballot(kind == 0), mb = __ballot(kind == 1), ml = __ballot(kind == 2);
        if (lane == 0) {
            off_a[k][wave] = (uint32_t)__popcll(ma);
            off_b[k][wave] = (uint32_t)__popcll(mb);
            off_l[k][wave] = (uint32_t)__popcll(ml);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t(*off)[4] = threadIdx.x == 0 ? off_a : (threadIdx.x == 1 ? off_b : off_l);
        uint32_t s = 0;
        for (int k = 0; k < GATE_PASSES; ++k)
            for (int w = 0; w < 4; ++w) {
                uint32_t t = off[k][w];
                off[k][w] = s;
                s += t;
            }
        base[threadIdx.x] = s ? atomicAdd(&counts[threadIdx.x], s) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GATE_PASSES; ++k) {
        int64_t id = chunk + k * 256 + threadIdx.x;
        uint32_t kind = (kinds >> (2 * k)) & 3u;
        unsigned long long ma = __ballot(kind == 0), mb = __ballot(kind == 1), ml = __ballot(kind == 2);
        if (kind == 0) list_a[base[0] + off_a[k][wave] + __popcll(ma & below)] = (int32_t)id;
        else if (kind == 1) list_b[base[1] + off_b[k][wave] + __popcll(mb & below)] = (int32_t)id;
        else if (kind == 2) list_l[base[2] + off_l[k][wave] + __popcll(ml & below)] = (int32_t)id;
    }
}

hipError_t launch_gate(const CamConst* cams, int n_cam, const float* tabs, int64_t n, double band, double crit_band,
                       double long_band, int32_t* list_a, int32_t* list_b, int32_t* list_l, uint32_t* counts,
                       uint8_t* flags, hipStream_t st) {
    hipLaunchKernelGGL(k_gate, dim3((unsigned)((n + GATE_RAYS - 1) / GATE_RAYS)), dim3(256), 0, st, cams, n_cam, tabs,
                       n, (float)band, (float)crit_band, (float)long_band, list_a, list_b, list_l, counts, flags);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Framebuffer shading of launch: linear RGB radiance per ray (RenderSession::shadeDiskHit,
// SNRS001A.cpp:315-434, for the thin disk's first crossing), then the average of each pixel's
// g x g rays (shadePixel, SNRS001A.cpp:480-560).
// ------------------------------------------------------------------------------------------
namespace {
// c00 w00 + c10 w10 + c01 w01 + c11 w11 on texel values / 255, each product and sum rounded on its own
__device__ __forceinline__ float bilinear(uint8_t a, uint8_t b, uint8_t c, uint8_t d, float w00, float w10, float w01,
                                          float w11) {
#pragma clang fp contract(off)
    // the IEEE float quotient (this file's FP32 division is the approximate one): in FP64, rounded
    // once, which is the correctly rounded float for every byte value
    const float fa = (float)(a / 255.0), fb = (float)(b / 255.0), fc = (float)(c / 255.0), fd = (float)(d / 255.0);
    return fa * w00 + fb * w10 + fc * w01 + fd * w11;
}

// CIE 1931 colour-matching functions by the piecewise-Gaussian fit of Wyman, Sloan and Shirley
// (JCGT 2(2), 2013): each lobe a exp(-((l - mu) s)^2 / 2), s taking one value below mu and another
// above it.
__device__ __forceinline__ double cie_lobe(double nm, double a, double mu, double s_lo, double s_hi) {
    const double t = (nm - mu) * (nm < mu ? s_lo : s_hi);
    return a * exp(-0.5 * t * t);
}

// Blackbody colour of temperature T (K), brightest channel 1 (PHSP001A.h:181-205): Planck's
// spectral radiance B(l, T) = c1 / (l^5 (e^{c2 / (l T)} - 1)) (c1 = 2 h c^2, c2 = h c / k, SI
// values; zero past c2 / (l T) = 700) weighs the matching functions at the 32 midpoints of
// 380-780 nm; the XYZ sums go through the sRGB (D65) matrix. The spectrum is FP64 and the colour
// sums float, each operation rounded on its own, as PHSP001A.h:72-80, 144-166 round them.
__device__ void blackbody_rgb(double T, float rgb[3]) {
#pragma clang fp contract(off)
    const double h = 6.62607015e-34, c = 2.99792458e8, k = 1.380649e-23;
    const double c1 = 2.0 * h * c * c, c2 = h * c / k;
    const double l0 = 380e-9, dl = (780e-9 - 380e-9) / 32;
    float X = 0.f, Y = 0.f, Z = 0.f;
    for (int i = 0; i < 32; ++i) {
        const double l = l0 + (i + 0.5) * dl, nm = l * 1e9;
        const double x = c2 / (l * T);
        const double B = (T <= 0 || x > 700) ? 0.0 : c1 / (pow(l, 5.0) * (exp(x) - 1.0));
        const float w = (float)(B * dl);
        const float sx = (float)(cie_lobe(nm, 0.362, 442.0, 0.0624, 0.0374) + cie_lobe(nm, 1.056, 599.8, 0.0264, 0.0323) -
                                 cie_lobe(nm, 0.065, 501.1, 0.0490, 0.0382));
        const float sy = (float)(cie_lobe(nm, 0.821, 568.8, 0.0213, 0.0247) + cie_lobe(nm, 0.286, 530.9, 0.0613, 0.0322));
        const float sz = (float)(cie_lobe(nm, 1.217, 437.0, 0.0845, 0.0278) + cie_lobe(nm, 0.681, 459.0, 0.0385, 0.0725));
        X = X + sx * w;
        Y = Y + sy * w;
        Z = Z + sz * w;
    }
    const float r = 3.2404542f * X - 1.5371385f * Y - 0.4985314f * Z;
    const float g = -0.9692660f * X + 1.8760108f * Y + 0.0415560f * Z;
    const float b = 0.0556434f * X - 0.2040259f * Y + 1.0572252f * Z;
    const float m = fmaxf(fmaxf(fmaxf(r, g), b), 0.001f);
    rgb[0] = __fdiv_rn(r, m);
    rgb[1] = __fdiv_rn(g, m);
    rgb[2] = __fdiv_rn(b, m);
}

// Linear limb-darkening law I(i) / I(0) = (1 + u cos i) / (1 + u) with u(l) falling linearly
// from 0.9 at 400 nm to 0.4 at 700 nm, at 650, 550 and 450 nm for red, green and blue
// (PHSP001A.h:280-322).
__device__ __forceinline__ float limb_factor(double lambda_nm, double cos_i) {
    const double t = fmin(fmax((lambda_nm - 400.0) / 300.0, 0.0), 1.0);
    const double u = 0.9 + t * (0.4 - 0.9);
    return (float)((1.0 + u * cos_i) / (1.0 + u));
}
}  // namespace

// One ray's radiance into fb[4 id..] (RGBA, alpha 1).
__global__ __launch_bounds__(256) void k_shade(const CamConst* __restrict__ cam, DevOut o, int64_t n, int emission,
                                               ShadeParams sp, const uint8_t* __restrict__ bg, int bgw, int bgh,
                                               float* __restrict__ fb) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    float rgb[3] = {0.f, 0.f, 0.f};
    const uint8_t cls = o.cls[id];
    if (cls == SIRIUS_RAY_DISK && emission == SIRIUS_EMISSION_THIN_DISK && cam->family == 0) {
#pragma clang fp contract(off)
        const CamConst& c = *cam;
        const double r = o.disk_r[id];
        // emitted temperature (K): Novikov-Thorne from the Page-Thorne flux, or Shakura-Sunyaev
        double te = 0.0;
        if (sp.temperature_model == 1) {
            const double f = page_thorne_f(c, r) / r;
            te = (c.fmax > 0 && f > 0) ? sp.disk_temperature * sqrt(sqrt(f / c.fmax)) : 0.0;
        } else {
            te = r > 0 ? sp.disk_temperature * pow(c.r_in / r, 0.75) : 0.0;
        }
        const float g = o.redshift[id];
        if (te > 0 && g > 0.f) {
            // direction of the photon at the hit: cos i = |k_z| / |k| (the disk's normal is z)
            const double* st = o.state + 8 * id;
            const double xd[4] = {st[0], st[1], st[2], st[3]}, pd[4] = {st[4], st[5], st[6], st[7]};
            const Spacetime<double> sd = make_spacetime<double>(c);
            double kd[4];
            raise_index<double, 0>(sd, xd, pd, kd);
            const double kn = sqrt(kd[1] * kd[1] + kd[2] * kd[2] + kd[3] * kd[3]);
            const double cos_i = kn > 0 ? fabs(kd[3]) / kn : 0.0;
            // TrueColor (PHSC001A.h:352-366): T_obs = T_e g in units of 30000 K, colour of
            // clamp(T_obs, 1000 K, 1e5 K), times T_obs^4 g^4
            const float t_obs = (float)(te / 30000.0) * g;
            const float tk = fminf(fmaxf(t_obs * 30000.0f, 1000.0f), 100000.0f);
            blackbody_rgb((double)tk, rgb);
            const float inten = t_obs * t_obs * t_obs * t_obs;
            const float g4 = g * g * g * g;
            const float s1 = inten * g4;
            const float limb[3] = {limb_factor(650.0, cos_i), limb_factor(550.0, cos_i), limb_factor(450.0, cos_i)};
            if (cos_i <= 0.0) rgb[0] = rgb[1] = rgb[2] = 0.f;
            for (int ch = 0; ch < 3; ++ch) rgb[ch] = ((rgb[ch] * s1) * limb[ch]) * g4 * (float)sp.disk_boost;
        }
    } else if (cls == SIRIUS_RAY_ESCAPED && bg) {
        // RenderSession::sampleStarfield (SNRS001A.cpp:631-703): phi in [0, 2 pi), u = phi / 2 pi,
        // v = theta / pi, bilinear between the four texels around (u (W - 1), v (H - 1)); the
        // coordinates in FP64 and the weights and sums in float, in the reference's order
        const double TWO_PI = 6.283185307179586, PI = 3.141592653589793;
        double ph = o.phi[id];
        if (ph < 0) ph += TWO_PI;
        const double u = ph / TWO_PI, v = (double)o.theta[id] / PI;
        const double fxp = u * (bgw - 1), fyp = v * (bgh - 1);
        const int x0 = min(bgw - 1, max(0, (int)floor(fxp))), y0 = min(bgh - 1, max(0, (int)floor(fyp)));
        const int x1 = min(x0 + 1, bgw - 1), y1 = min(y0 + 1, bgh - 1);
        const double fx = fxp - x0, fy = fyp - y0;
        const float w00 = (float)((1.0 - fx) * (1.0 - fy)), w10 = (float)(fx * (1.0 - fy));
        const float w01 = (float)((1.0 - fx) * fy), w11 = (float)(fx * fy);
        const uint8_t* c00 = bg + 4 * ((int64_t)y0 * bgw + x0);
        const uint8_t* c10 = bg + 4 * ((int64_t)y0 * bgw + x1);
        const uint8_t* c01 = bg + 4 * ((int64_t)y1 * bgw + x0);
        const uint8_t* c11 = bg + 4 * ((int64_t)y1 * bgw + x1);
        for (int ch = 0; ch < 3; ++ch) rgb[ch] = bilinear(c00[ch], c10[ch], c01[ch], c11[ch], w00, w10, w01, w11);
    } else if (cls == SIRIUS_RAY_MAX_STEPS || cls == SIRIUS_RAY_FAILED) {
        rgb[0] = rgb[1] = rgb[2] = (float)sp.max_steps_level;
    }
    fb[4 * id + 0] = rgb[0];
    fb[4 * id + 1] = rgb[1];
    fb[4 * id + 2] = rgb[2];
    fb[4 * id + 3] = 1.f;
}

// Average of each pixel's g x g rays (sub-image of gW x gH pixels, row-major) into the W x H frame.
__global__ __launch_bounds__(256) void k_pixel_average(const float* __restrict__ sub, int w, int h, int gs,
                                                       float* __restrict__ fb) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)w * h) return;
    const int px = (int)(id % w), py = (int)(id / w);
    const int64_t sw = (int64_t)w * gs;
    float acc[3] = {0.f, 0.f, 0.f};
    for (int sy = 0; sy < gs; ++sy)
        for (int sx = 0; sx < gs; ++sx) {
            const int64_t s = ((int64_t)py * gs + sy) * sw + (int64_t)px * gs + sx;
            for (int ch = 0; ch < 3; ++ch) acc[ch] += sub[4 * s + ch];
        }
    const float inv = 1.f / (float)(gs * gs);
    for (int ch = 0; ch < 3; ++ch) fb[4 * id + ch] = acc[ch] * inv;
    fb[4 * id + 3] = 1.f;
}

hipError_t launch_shade(const CamConst* cam, const DevOut& o, int64_t n, int emission, const ShadeParams& sp,
                        const uint8_t* bg, int bgw, int bgh, float* fb, hipStream_t st) {
    hipLaunchKernelGGL(k_shade, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cam, o, n, emission, sp, bg, bgw,
                       bgh, fb);
    return hipGetLastError();
}

hipError_t launch_pixel_average(const float* sub, int w, int h, int gs, float* fb, hipStream_t st) {
    const int64_t n = (int64_t)w * h;
    hipLaunchKernelGGL(k_pixel_average, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sub, w, h, gs, fb);
    return hipGetLastError();
}


// ------------------------------------------------------------------------------------------
// Celestial spheres: StarfieldGenerator::sampleStarfield (PHSF001A.h:153-201) for an observer at
// the origin. Star records (SIRIUS_STAR_REC floats, see starfield.cpp) are staged through LDS in
// tiles of 256; each lane sums the Gaussian point-spread of every star of the tile whose
// direction lies within its cosine cut. The view-dependent float steps are the reference's, in
// its order and unfused, so the cosine a pixel sees - which decides whether a point-like star
// lights it - matches the reference's bit for bit.
// ------------------------------------------------------------------------------------------
namespace {
constexpr int STAR_TILE = 256;

__device__ __forceinline__ void star_light(const float* __restrict__ tile, int cnt, float brightness, float vx,
                                           float vy, float vz, float& r, float& g, float& b) {
    // Every product and sum rounds on its own: contraction into fused multiply-adds is off in
    // this scope (HIP's __fmul_rn / __fadd_rn would not prevent it: their bodies are plain
    // operators that still fuse once inlined), and the division is the IEEE one.
#pragma clang fp contract(off)
    for (int j = 0; j < cnt; ++j) {
        const float* s = tile + j * SIRIUS_STAR_REC;
        float c = (vx * s[0] + vy * s[1]) + vz * s[2];
        if (c < s[3]) continue;
        // acos and exp correctly rounded to float (FP64, then rounded): near a cosine of 1 the
        // angle jumps by ~3.5e-4 rad per float step, so the spread of a point-like star depends
        // on every bit of it; the C library's float acos is correctly rounded there too
        float ang = (float)acos((double)fminf(fmaxf(c, -1.0f), 1.0f));
        float psf = (float)exp((double)__fdiv_rn((-0.5f * ang) * ang, s[4]));
        if (psf < 1e-6f) continue;
        float inten = (s[5] * psf) * brightness;
        r = r + s[6] * inten;
        g = g + s[7] * inten;
        b = b + s[8] * inten;
    }
}

// Block-wide pass over one catalogue; every thread of the block takes part in the staging,
// lanes with want == false skip the sums.
__device__ __forceinline__ void star_pass(float* tile, const float* __restrict__ rec, int n_star, float brightness,
                                          bool want, float vx, float vy, float vz, float& r, float& g, float& b) {
    for (int base = 0; base < n_star; base += STAR_TILE) {
        int cnt = min(STAR_TILE, n_star - base);
        __syncthreads();
        for (int k = threadIdx.x; k < cnt * SIRIUS_STAR_REC; k += blockDim.x)
            tile[k] = rec[(int64_t)base * SIRIUS_STAR_REC + k];
        __syncthreads();
        if (want) star_light(tile, cnt, brightness, vx, vy, vz, r, g, b);
    }
}
}  // namespace

__global__ __launch_bounds__(256) void k_star_dirs(const float* __restrict__ rec, int n_star, float brightness,
                                                   const float* __restrict__ dirs, int n, float* __restrict__ rgb) {
    __shared__ float tile[STAR_TILE * SIRIUS_STAR_REC];
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    bool act = id < n;
    float vx = act ? dirs[3 * id] : 0.f, vy = act ? dirs[3 * id + 1] : 0.f, vz = act ? dirs[3 * id + 2] : 0.f;
    float r = 0.f, g = 0.f, b = 0.f;
    star_pass(tile, rec, n_star, brightness, act, vx, vy, vz, r, g, b);
    if (act) {
        rgb[3 * id] = r;
        rgb[3 * id + 1] = g;
        rgb[3 * id + 2] = b;
    }
}

// Star light added to the framebuffer of escaped rays: the sky (or a wormhole's near side) from
// sphere 0, a wormhole's far side (SIRIUS_FLAG_FAR_SIDE) from sphere 1; the direction is the
// ray's escape direction (theta, phi).
// A pixel's beam on the sky: the footprint covariance c00, c01, c11 (rad^2) in the frame
// (e_theta, e_phi) at the escape direction, and the cut's widening (cos, sin of its extent).
struct Beam {
    float eth[3], eph[3];
    float c00, c01, c11;
    float cb, sb;
};

// Star light averaged over a beam: the reference's point spread exp(-a^2 / (2 sigma^2)) (peak 1)
// convolved with the beam's Gaussian footprint, evaluated in the tangent plane at the pixel's
// direction: exp(-d^T S^-1 d / 2) sigma^2 / sqrt(det S), S = C + sigma^2 I. With a vanishing
// footprint it is the point-sampled light; summed over pixels it keeps each star's integrated
// light, which point sampling catches only where a pixel centre falls inside the spread.
__device__ __forceinline__ void star_light_beam(const float* tile, int cnt, float brightness, float vx, float vy,
                                                float vz, const Beam& bm, float& r, float& g, float& b) {
    for (int j = 0; j < cnt; ++j) {
        const float* s = tile + j * SIRIUS_STAR_REC;
        const float c = vx * s[0] + vy * s[1] + vz * s[2];
        // within the star's cut angle plus the beam's extent: cos(A + B) = cos A cos B - sin A sin B
        if (c < s[3] * bm.cb - s[9] * bm.sb) continue;
        const float d0 = s[0] * bm.eth[0] + s[1] * bm.eth[1] + s[2] * bm.eth[2];
        const float d1 = s[0] * bm.eph[0] + s[1] * bm.eph[1] + s[2] * bm.eph[2];
        const float s00 = bm.c00 + s[4], s11 = bm.c11 + s[4], s01 = bm.c01;
        const float det = s00 * s11 - s01 * s01;
        const float q = (s11 * d0 * d0 - 2.0f * s01 * d0 * d1 + s00 * d1 * d1) / det;
        const float shape = expf(-0.5f * q);
        if (shape < 1e-6f) continue;
        const float inten = s[5] * shape * (s[4] / sqrtf(det)) * brightness;
        r += s[6] * inten;
        g += s[7] * inten;
        b += s[8] * inten;
    }
}

// One walk over a catalogue's tiles; each lane point-samples or, with `beam`, averages over its beam.
__device__ __forceinline__ void star_pass_mixed(float* tile, const float* __restrict__ rec, int n_star,
                                                float brightness, bool want, bool beam, float vx, float vy, float vz,
                                                const Beam& bm, float& r, float& g, float& b) {
    for (int base = 0; base < n_star; base += STAR_TILE) {
        int cnt = min(STAR_TILE, n_star - base);
        __syncthreads();
        for (int k = threadIdx.x; k < cnt * SIRIUS_STAR_REC; k += blockDim.x)
            tile[k] = rec[(int64_t)base * SIRIUS_STAR_REC + k];
        __syncthreads();
        if (want) {
            if (beam) star_light_beam(tile, cnt, brightness, vx, vy, vz, bm, r, g, b);
            else star_light(tile, cnt, brightness, vx, vy, vz, r, g, b);
        }
    }
}

// Star light of escaped pixels. With a bundle (bjac non-null), a pixel whose beam escaped and
// was not terminated gets the light averaged over its beam (radius bsize pixels at the camera);
// the others, and every pixel without a bundle, are point-sampled.
__global__ __launch_bounds__(256) void k_star_shade(DevOut o, int64_t n, const float* __restrict__ rec0, int n0,
                                                    float b0, const float* __restrict__ rec1, int n1, float b1,
                                                    const float* __restrict__ bjac, const uint8_t* __restrict__ bcls,
                                                    const uint8_t* __restrict__ bflags, float bsize,
                                                    float* __restrict__ fb) {
    __shared__ float tile[STAR_TILE * SIRIUS_STAR_REC];
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool esc = id < n && o.cls[id] == SIRIUS_RAY_ESCAPED;
    bool far = esc && (o.flags[id] & SIRIUS_FLAG_FAR_SIDE);
    float vx = 0.f, vy = 0.f, vz = 0.f;
    Beam bm;
    bool beam = false;
    if (esc) {
        // the unit escape direction in FP64, each component rounded once to float: a host that
        // forms it the same way gets the same bits (to a double rounding near a float midpoint),
        // and next to a point-like star one ulp of the direction decides whether it lights
        const double th = o.theta[id], ph = o.phi[id];
        const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
        vx = (float)(st * cp);
        vy = (float)(st * sp);
        vz = (float)ct;
        if (bjac && bcls[id] == SIRIUS_RAY_ESCAPED && !(bflags[id] & SIRIUS_BUNDLE_TERMINATED)) {
            // footprint covariance s^2 J J^T with J's rows d theta/d(u, v), sin(theta) d phi/d(u, v)
            const double j00 = bjac[id], j01 = bjac[n + id];
            const double j10 = st * bjac[2 * n + id], j11 = st * bjac[3 * n + id];
            const double s2 = (double)bsize * bsize;
            const double c00 = s2 * (j00 * j00 + j01 * j01), c11 = s2 * (j10 * j10 + j11 * j11);
            const double c01 = s2 * (j00 * j10 + j01 * j11);
            const double tr = c00 + c11, det = c00 * c11 - c01 * c01;
            const double lmax = 0.5 * (tr + sqrt(fmax(0.0, tr * tr - 4.0 * det)));
            const double ext = 1.01 * sqrt(2.0 * log(1e6)) * sqrt(lmax);
            if (ext < 0.5) {  // a tangent-plane beam: small against a radian
                beam = true;
                bm.eth[0] = (float)(ct * cp); bm.eth[1] = (float)(ct * sp); bm.eth[2] = (float)(-st);
                bm.eph[0] = (float)(-sp); bm.eph[1] = (float)cp; bm.eph[2] = 0.f;
                bm.c00 = (float)c00; bm.c01 = (float)c01; bm.c11 = (float)c11;
                bm.cb = (float)cos(ext);
                bm.sb = (float)sin(ext);
            }
        }
    }
    float r = 0.f, g = 0.f, b = 0.f;
    if (bjac) {
        star_pass_mixed(tile, rec0, n0, b0, esc && !far, beam, vx, vy, vz, bm, r, g, b);
        star_pass_mixed(tile, rec1, n1, b1, far, beam, vx, vy, vz, bm, r, g, b);
    } else {
        star_pass(tile, rec0, n0, b0, esc && !far, vx, vy, vz, r, g, b);
        star_pass(tile, rec1, n1, b1, far, vx, vy, vz, r, g, b);
    }
    if (esc) {
        fb[4 * id + 0] += r;
        fb[4 * id + 1] += g;
        fb[4 * id + 2] += b;
    }
}

hipError_t launch_star_dirs(const float* rec, int n_star, float brightness, const float* dirs, int n, float* rgb,
                            hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_star_dirs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rec, n_star, brightness, dirs,
                       n, rgb);
    return hipGetLastError();
}

hipError_t launch_star_shade(const DevOut& o, int64_t n, const float* rec0, int n0, float b0, const float* rec1, int n1,
                             float b1, const float* bjac, const uint8_t* bcls, const uint8_t* bflags, float bsize,
                             float* fb, hipStream_t st) {
    if (n <= 0 || (n0 <= 0 && n1 <= 0)) return hipSuccess;
    hipLaunchKernelGGL(k_star_shade, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, o, n, rec0, n0, b0, rec1, n1,
                       b1, bjac, bcls, bflags, bsize, fb);
    return hipGetLastError();
}

}  // namespace sirius
