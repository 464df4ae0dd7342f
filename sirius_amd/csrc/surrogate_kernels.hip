// surrogate_kernels.hip — per-ray kernels of the far-field surrogate: network inputs and
// targets, the accept / fall-back decision with the work lists it feeds, and the audit.
#include <hip/hip_runtime.h>

#include "geodesic.hpp"
#include "mlp.hpp"

namespace sirius {

__device__ __forceinline__ int sur_camera(const CamConst* __restrict__ cams, int n_cam, int64_t id) {
    int lo = 0, hi = n_cam - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cams[mid].offset <= id) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Network inputs (a, theta_obs / (pi/2), (r_obs - 60) / 40, alpha / 15, beta / 15), the
// initial Cartesian direction n0, and |distance| from the critical curve (tables: tabs, or
// +inf without them) of every ray; with flags, every ray's flag byte is initialised (the
// critical-band flag when |d| < crit_band).
__global__ __launch_bounds__(256) void k_sur_features(const CamConst* __restrict__ cams, int n_cam,
                                                      const float* __restrict__ tabs, int64_t n, float crit_band,
                                                      float* __restrict__ x, float* __restrict__ n0,
                                                      float* __restrict__ dist, uint8_t* __restrict__ flags) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int cam = sur_camera(cams, n_cam, id);
    const CamConst& c = cams[cam];
    const int64_t pix = id - c.offset;
    const int px = (int)(pix % c.width), py = (int)(pix / c.width);
    Spacetime<double> s = make_spacetime<double>(c);
    double xs[4], ps[4];
    init_state<double>(c, s, px, py, xs, ps);
    double alpha, beta;
    sky_alpha_beta(s, xs, ps, alpha, beta);
    x[id * 5 + 0] = (float)c.a;
    x[id * 5 + 1] = (float)(c.theta_obs / 1.5707963267948966);
    x[id * 5 + 2] = (float)((c.r_obs - 60.0) / 40.0);
    x[id * 5 + 3] = (float)(alpha / 15.0);
    x[id * 5 + 4] = (float)(beta / 15.0);
    double nd[3];
    camera_cart_dir<double>(c, px, py, nd);
    for (int k = 0; k < 3; ++k) n0[id * 3 + k] = (float)nd[k];
    float d = INFINITY;
    if (tabs) d = fabsf((float)crit_distance<double>(tabs + (int64_t)cam * (SIRIUS_CC_TABLE + 1), s, xs, ps));
    if (dist) dist[id] = d;
    if (flags) flags[id] = d < crit_band ? (uint8_t)SIRIUS_FLAG_CRITICAL : (uint8_t)0;
}

// Training targets from integrator results: class 0 escaped, 2 disk, 1 anything else; the
// escape direction residual d_esc - n0 for escaped rays.
__global__ void k_sur_targets(DevOut o, const float* __restrict__ n0, int64_t n, int32_t* __restrict__ cls,
                              float* __restrict__ dir) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int c = o.cls[id];
    cls[id] = c == SIRIUS_RAY_ESCAPED ? 0 : (c == SIRIUS_RAY_DISK ? 2 : 1);
    float d[3] = {0.f, 0.f, 0.f};
    if (c == SIRIUS_RAY_ESCAPED) {
        float th = o.theta[id], ph = o.phi[id];
        d[0] = sinf(th) * cosf(ph) - n0[id * 3 + 0];
        d[1] = sinf(th) * sinf(ph) - n0[id * 3 + 1];
        d[2] = cosf(th) - n0[id * 3 + 2];
    }
    for (int k = 0; k < 3; ++k) dir[id * 3 + k] = d[k];
}

__device__ __forceinline__ uint64_t sur_hash(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Decision per ray. Answered (written here, SIRIUS_FLAG_SURROGATE): outside the fallback band,
// the most probable class escaped or captured with probability >= pmin. Every other ray goes
// to the integrator: list B (FP64 band, |d| < esc_band) or list A (FP32). A seeded 1-in-
// (1 / audit_fraction) subset of answered rays also goes to the audit list. Lists are filled
// with one atomic per block and list (ray order kept within a block).
constexpr int SUR_PASSES = 8;
constexpr int SUR_RAYS = 256 * SUR_PASSES;

__global__ __launch_bounds__(256) void k_sur_apply(const float* __restrict__ y, const float* __restrict__ n0,
                                                   const float* __restrict__ dist, int64_t n, float esc_band,
                                                   float sur_band, float pmin, double audit_fraction,
                                                   uint64_t audit_seed, DevOut o, int n_bins,
                                                   int32_t* __restrict__ list_a, int32_t* __restrict__ list_b,
                                                   int32_t* __restrict__ list_audit, uint32_t* __restrict__ counts) {
    __shared__ uint32_t off[3][SUR_PASSES][4];
    __shared__ uint32_t base[3];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const int64_t chunk = (int64_t)blockIdx.x * SUR_RAYS;
    const unsigned long long below = lane ? ((1ull << lane) - 1ull) : 0ull;
    uint32_t kind_bits = 0;  // 2 bits per pass: 0 answered, 1 list A, 2 list B, 3 invalid
    uint32_t audit_bits = 0;
    const uint64_t thr = audit_fraction >= 1.0 ? ~0ull : (uint64_t)(audit_fraction * 18446744073709551615.0);
#pragma unroll
    for (int k = 0; k < SUR_PASSES; ++k) {
        const int64_t id = chunk + k * 256 + threadIdx.x;
        uint32_t kind = 3;
        bool audit = false;
        if (id < n) {
            const float* yy = y + id * SIRIUS_MLP_OUT;
            float m = fmaxf(yy[0], fmaxf(yy[1], yy[2]));
            float e0 = expf(yy[0] - m), e1 = expf(yy[1] - m), e2 = expf(yy[2] - m);
            float s = e0 + e1 + e2;
            int arg = (e0 >= e1 && e0 >= e2) ? 0 : (e1 >= e2 ? 1 : 2);
            float pmax = fmaxf(e0, fmaxf(e1, e2)) / s;
            float d = dist[id];
            bool ok = d >= sur_band && pmax >= pmin && arg != 2;
            if (ok) {
                kind = 0;
                uint8_t fl = o.flags[id] | (uint8_t)SIRIUS_FLAG_SURROGATE;
                audit = audit_fraction > 0.0 && sur_hash(audit_seed ^ (uint64_t)id) <= thr;
                if (audit) fl |= (uint8_t)SIRIUS_FLAG_AUDITED;
                o.flags[id] = fl;
                o.cls[id] = arg == 0 ? (uint8_t)SIRIUS_RAY_ESCAPED : (uint8_t)SIRIUS_RAY_CAPTURED;
                float th = 0.f, ph = 0.f;
                if (arg == 0) {
                    float v0 = n0[id * 3 + 0] + yy[3], v1 = n0[id * 3 + 1] + yy[4], v2 = n0[id * 3 + 2] + yy[5];
                    float il = 1.f / sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
                    th = acosf(fminf(1.f, fmaxf(-1.f, v2 * il)));
                    ph = atan2f(v1, v0);
                }
                o.theta[id] = th;
                o.phi[id] = ph;
                o.lam[id] = 0.f;
                o.disk_r[id] = 0.f;
                o.redshift[id] = 0.f;
                o.null_drift[id] = 0.f;
                o.steps[id] = 0u;
                o.rejected[id] = 0u;
                o.landings[id] = 0u;
                for (int b = 0; b < n_bins; ++b) o.intensity[(int64_t)b * o.n_rays + id] = 0.f;
            } else {
                kind = d < esc_band ? 2u : 1u;
            }
        }
        kind_bits |= kind << (2 * k);
        audit_bits |= (uint32_t)audit << k;
        unsigned long long ma = __ballot(kind == 1), mb = __ballot(kind == 2), mu = __ballot(audit);
        if (lane == 0) {
            off[0][k][wave] = (uint32_t)__popcll(ma);
            off[1][k][wave] = (uint32_t)__popcll(mb);
            off[2][k][wave] = (uint32_t)__popcll(mu);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int L = threadIdx.x;
        uint32_t s = 0;
        for (int k = 0; k < SUR_PASSES; ++k)
            for (int w = 0; w < 4; ++w) {
                uint32_t t = off[L][k][w];
                off[L][k][w] = s;
                s += t;
            }
        base[L] = s ? atomicAdd(&counts[L], s) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SUR_PASSES; ++k) {
        const int64_t id = chunk + k * 256 + threadIdx.x;
        const uint32_t kind = (kind_bits >> (2 * k)) & 3u;
        const bool audit = (audit_bits >> k) & 1u;
        unsigned long long ma = __ballot(kind == 1), mb = __ballot(kind == 2), mu = __ballot(audit);
        if (kind == 1) list_a[base[0] + off[0][k][wave] + __popcll(ma & below)] = (int32_t)id;
        if (kind == 2) list_b[base[1] + off[1][k][wave] + __popcll(mb & below)] = (int32_t)id;
        if (audit) list_audit[base[2] + off[2][k][wave] + __popcll(mu & below)] = (int32_t)id;
    }
}

// Training batch gather through a permutation.
__global__ void k_sur_gather(const int32_t* __restrict__ perm, int64_t start, int64_t m, const float* __restrict__ x,
                             const int32_t* __restrict__ cls, const float* __restrict__ dir, float* __restrict__ xb,
                             int32_t* __restrict__ cb, float* __restrict__ db) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t j = perm[start + i];
    for (int k = 0; k < 5; ++k) xb[i * 5 + k] = x[j * 5 + k];
    for (int k = 0; k < 3; ++k) db[i * 3 + k] = dir[j * 3 + k];
    cb[i] = cls[j];
}
hipError_t launch_sur_gather(const int32_t* perm, int64_t start, int64_t m, const float* x, const int32_t* cls,
                             const float* dir, float* xb, int32_t* cb, float* db, hipStream_t st) {
    hipLaunchKernelGGL(k_sur_gather, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, perm, start, m, x, cls, dir,
                       xb, cb, db);
    return hipGetLastError();
}

hipError_t launch_sur_features(const CamConst* cams, int n_cam, const float* tabs, int64_t n, double crit_band,
                               float* x, float* n0, float* dist, uint8_t* flags, hipStream_t st) {
    hipLaunchKernelGGL(k_sur_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cams, n_cam, tabs, n,
                       (float)crit_band, x, n0, dist, flags);
    return hipGetLastError();
}
hipError_t launch_sur_targets(const DevOut& o, const float* n0, int64_t n, int32_t* cls, float* dir, hipStream_t st) {
    hipLaunchKernelGGL(k_sur_targets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, o, n0, n, cls, dir);
    return hipGetLastError();
}
hipError_t launch_sur_apply(const float* y, const float* n0, const float* dist, int64_t n, double esc_band,
                            double sur_band, double pmin, double audit_fraction, uint64_t audit_seed, const DevOut& o,
                            int n_bins, int32_t* list_a, int32_t* list_b, int32_t* list_audit, uint32_t* counts,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_sur_apply, dim3((unsigned)((n + SUR_RAYS - 1) / SUR_RAYS)), dim3(256), 0, st, y, n0, dist, n,
                       (float)esc_band, (float)sur_band, (float)pmin, audit_fraction, audit_seed, o, n_bins, list_a,
                       list_b, list_audit, counts);
    return hipGetLastError();
}
}  // namespace sirius
