// surrogate_kernels.hip — per-ray kernels of the far-field surrogate: the analytic candidate test
// and reference direction, network inputs and training targets, the gate that sorts a batch into
// the integrator's lists and the network's candidates, the accept / fall-back decision, the
// held-out calibration of the error gate, and the audit's comparison.
#include <hip/hip_runtime.h>

#include "farfield.hpp"
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

// Bin of the calibrated error gate: spin (10 bins of 0.1), distance from the critical curve
// (edges 0.5 1 1.5 2 3 4 6 8 12 M) and observer radius (20-46.7-73.3-100 M).
__device__ __forceinline__ int sur_bin(float a, float d, float r_obs) {
    int ia = (int)(fabsf(a) * 10.f);
    ia = ia < 0 ? 0 : (ia > 9 ? 9 : ia);
    const float e[8] = {1.f, 1.5f, 2.f, 3.f, 4.f, 6.f, 8.f, 12.f};
    int id = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) id += d >= e[k] ? 1 : 0;
    const int ir = r_obs < 46.666667f ? 0 : (r_obs < 73.333333f ? 1 : 2);
    return (ia * 9 + id) * 3 + ir;
}

// The bin of a ray from its record (spin rec[0], distance rec[5], scaled observer radius rec[2]).
__device__ __forceinline__ int sur_bin_rec(const float* rr) { return sur_bin(rr[0], rr[5], rr[2] * 40.f + 60.f); }

// Calibration of the error gate: a bin answers when its held-out rays number at least
// SUR_CAL_MIN, none was misclassified, and the largest direction error is within err_max; it
// answers a ray only within the distance from the critical curve that its held-out answers
// reached (no extrapolation beyond the evidence, such as the corners of a wider frame).
struct SurCal {
    const float* err;
    const uint32_t* n;
    const uint32_t* bad;
    const float* dmax;
    float err_max;
    float sup_x, sup_y;  // the training cameras' image-plane box (see MlpState)
};
__device__ __forceinline__ bool sur_bin_usable(const SurCal& cal, int b, float d) {
    return cal.n[b] >= SUR_CAL_MIN && cal.bad[b] == 0u && cal.err[b] <= cal.err_max && d <= cal.dmax[b];
}

// The analytic reference of one ray (FP32, from the camera): its signed distance from the
// critical curve, whether the surrogate may be asked (outside the fallback band and the shadow,
// inside the training cameras' field of view and in a bin of `cal` that answers when cal is
// given, its Schwarzschild orbit scattered and crossing
// the equatorial plane only outside the thin disk widened by `margin`), and for such a candidate
// the record rec[SUR_REC] (network inputs, distance, b, reference exit direction). Returns the
// candidate decision.
__device__ __forceinline__ bool sur_reference(const CamConst& c, const float* __restrict__ tab, int px, int py,
                                              float sur_band, float margin, float r_exit, float& dist, float* rec,
                                              const SurCal* cal = nullptr) {
    Spacetime<float> s = make_spacetime<float>(c);
    float x[4], p[4];
    init_state<float>(c, s, px, py, x, p);
    dist = crit_distance<float>(tab, s, x, p);
    if (!(dist >= sur_band)) return false;
    if (cal) {
        // inside the training cameras' field of view (the same image-plane coordinates as the
        // camera's ray construction)
        const float sx = ((2.f * (px + 0.5f)) / (float)c.width - 1.f) * (float)c.aspect * (float)c.tan_half;
        const float sy = (1.f - (2.f * (py + 0.5f)) / (float)c.height) * (float)c.tan_half;
        if (!(fabsf(sx) <= cal->sup_x && fabsf(sy) <= cal->sup_y)) return false;
    }
    // a ray whose calibration bin cannot answer skips the orbit test and the network (the bin
    // from the same FP32 values the record carries)
    if (cal && !sur_bin_usable(*cal, sur_bin((float)c.a, dist, (float)((c.r_obs - 60.0) / 40.0) * 40.f + 60.f), dist))
        return false;
    float n[3];
    camera_cart_dir<float>(c, px, py, n);
    // to first order in a, frame dragging deflects a distant ray as if the lens sat a away from the
    // hole, along -e_phi of the observer; 1.1 a is the displacement that minimises the residual
    // over the training distribution (spins to 0.99)
    const float sh = (float)(SUR_SPIN_SHIFT * c.a);
    const float x0[3] = {x[1] + sh * (float)c.ep[0], x[2] + sh * (float)c.ep[1], x[3] + sh * (float)c.ep[2]};
    const FarOrbit<float> o8 = far_orbit<8, float>(x0, n, r_exit);
    if (!o8.ok) return false;
    if (far_crosses_annulus<8, float>(o8, (float)c.r_in - margin, (float)c.r_out + margin)) return false;
    if (rec) {
        float alpha, beta;
        sky_alpha_beta(s, x, p, alpha, beta);
        rec[0] = (float)c.a;
        rec[1] = (float)(c.theta_obs / 1.5707963267948966);
        rec[2] = (float)((c.r_obs - 60.0) / 40.0);
        rec[3] = alpha / 15.f;
        rec[4] = beta / 15.f;
        rec[5] = dist;
        const FarOrbit<float> o16 = far_orbit<16, float>(x0, n, r_exit);
        rec[6] = o16.b;
        far_exit_dir(o16, rec + 8);
        rec[7] = 0.f;
        rec[11] = 0.f;
    }
    return true;
}

// Network inputs and reference directions of every ray (the training set and the public
// features call): x[5 id..], ref[3 id..] (zero for rays the analytic test refuses), cand[id].
__global__ __launch_bounds__(256) void k_sur_features(const CamConst* __restrict__ cams, int n_cam,
                                                      const float* __restrict__ tabs, int64_t n, float sur_band,
                                                      float margin, float r_exit, float* __restrict__ x,
                                                      float* __restrict__ ref, uint8_t* __restrict__ cand,
                                                      float* __restrict__ dist_out, float* __restrict__ rec_out) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int cam = sur_camera(cams, n_cam, id);
    const CamConst& c = cams[cam];
    const int64_t pix = id - c.offset;
    const int px = (int)(pix % c.width), py = (int)(pix / c.width);
    float rec[SUR_REC];
    float d;
    const bool ok = c.family == 0 &&
                    sur_reference(c, tabs + (int64_t)cam * (SIRIUS_CC_TABLE + 1), px, py, sur_band, margin, r_exit, d, rec);
    if (!ok) {
        // inputs of a refused ray still come from its camera (FP32 initial state)
        Spacetime<float> s = make_spacetime<float>(c);
        float xs[4], ps[4], al = 0.f, be = 0.f;
        if (c.family == 0) {
            init_state<float>(c, s, px, py, xs, ps);
            sky_alpha_beta(s, xs, ps, al, be);
        }
        rec[0] = (float)c.a;
        rec[1] = (float)(c.theta_obs / 1.5707963267948966);
        rec[2] = (float)((c.r_obs - 60.0) / 40.0);
        rec[3] = al / 15.f;
        rec[4] = be / 15.f;
        rec[8] = rec[9] = rec[10] = 0.f;
    }
    for (int k = 0; k < SIRIUS_MLP_IN; ++k) x[id * SIRIUS_MLP_IN + k] = rec[k];
    for (int k = 0; k < 3; ++k) ref[id * 3 + k] = rec[8 + k];
    if (cand) cand[id] = ok ? 1 : 0;
    if (dist_out) dist_out[id] = ok ? d : 0.f;
    if (rec_out && ok)
        for (int k = 0; k < SUR_REC; ++k) rec_out[id * SUR_REC + k] = rec[k];
}

// Training targets from integrator results: class 0 escaped, 2 disk, 1 anything else; for
// escaped rays the residual of the escape direction over the reference, (d_esc - ref) / scale.
__global__ void k_sur_targets(DevOut o, const float* __restrict__ ref, int64_t n, int32_t* __restrict__ cls,
                              float* __restrict__ dir) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int c = o.cls[id];
    cls[id] = c == SIRIUS_RAY_ESCAPED ? 0 : (c == SIRIUS_RAY_DISK ? 2 : 1);
    float d[3] = {0.f, 0.f, 0.f};
    if (c == SIRIUS_RAY_ESCAPED) {
        const double th = o.theta[id], ph = o.phi[id];
        const double e[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
        for (int k = 0; k < 3; ++k) d[k] = (float)((e[k] - (double)ref[id * 3 + k]) / SIRIUS_MLP_DIR_SCALE);
    }
    for (int k = 0; k < 3; ++k) dir[id * 3 + k] = d[k];
}

__device__ __forceinline__ uint64_t sur_hash(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Lists by one atomic per block and list: `kinds` holds each lane's list for pass k (3 bits each),
// offsets are the per-wave counts, turned into the block's running offsets here.
constexpr int SG_PASSES = 8;
constexpr int SG_RAYS = 256 * SG_PASSES;

// Gate of the FP32 path with the surrogate: like the plain gate (flags, band list B, long list L,
// list A), with the network's candidates in list S; candidates leave their record in rec[id].
// Only rays in a bin whose calibration answers are candidates: with an uncalibrated network, or
// a scene whose bins the calibration refuses, the network does not run. counts[0..3] = |A|, |B|,
// |L|, |S|.
__global__ __launch_bounds__(256) void k_sur_gate(const CamConst* __restrict__ cams, int n_cam,
                                                  const float* __restrict__ tabs, int64_t n, float band, float crit_band,
                                                  float long_band, float sur_band, float margin, float r_exit,
                                                  int32_t* __restrict__ list_a, int32_t* __restrict__ list_b,
                                                  int32_t* __restrict__ list_l, int32_t* __restrict__ list_s,
                                                  uint32_t* __restrict__ counts, uint8_t* __restrict__ flags,
                                                  float* __restrict__ rec, SurCal cal) {
    __shared__ uint32_t off[4][SG_PASSES][4];
    __shared__ uint32_t base[4];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const int64_t chunk = (int64_t)blockIdx.x * SG_RAYS;
    const unsigned long long below = lane ? ((1ull << lane) - 1ull) : 0ull;
    uint32_t kinds = 0;  // 3 bits per pass: 0 A, 1 B, 2 L, 3 S, 4 invalid
#pragma unroll
    for (int k = 0; k < SG_PASSES; ++k) {
        const int64_t id = chunk + k * 256 + threadIdx.x;
        uint32_t kind = 4;
        if (id < n) {
            const int cam = sur_camera(cams, n_cam, id);
            const CamConst& c = cams[cam];
            const int64_t pix = id - c.offset;
            float d;
            const bool cand = sur_reference(c, tabs + (int64_t)cam * (SIRIUS_CC_TABLE + 1), (int)(pix % c.width),
                                            (int)(pix / c.width), sur_band, margin, r_exit, d, rec + id * SUR_REC,
                                            &cal);
            const float ad = fabsf(d);
            flags[id] = ad < crit_band ? (uint8_t)SIRIUS_FLAG_CRITICAL : (uint8_t)0;
            kind = ad < band ? 1u : (cand ? 3u : (ad < long_band ? 2u : 0u));
        }
        kinds |= kind << (3 * k);
#pragma unroll
        for (int L = 0; L < 4; ++L) {
            const unsigned long long m = __ballot(kind == (uint32_t)L);
            if (lane == 0) off[L][k][wave] = (uint32_t)__popcll(m);
        }
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int L = threadIdx.x;
        uint32_t s = 0;
        for (int k = 0; k < SG_PASSES; ++k)
            for (int w = 0; w < 4; ++w) {
                const uint32_t t = off[L][k][w];
                off[L][k][w] = s;
                s += t;
            }
        base[L] = s ? atomicAdd(&counts[L], s) : 0u;
    }
    __syncthreads();
    int32_t* lists[4] = {list_a, list_b, list_l, list_s};
#pragma unroll
    for (int k = 0; k < SG_PASSES; ++k) {
        const int64_t id = chunk + k * 256 + threadIdx.x;
        const uint32_t kind = (kinds >> (3 * k)) & 7u;
#pragma unroll
        for (int L = 0; L < 4; ++L) {
            const unsigned long long m = __ballot(kind == (uint32_t)L);
            if (kind == (uint32_t)L) lists[L][base[L] + off[L][k][wave] + __popcll(m & below)] = (int32_t)id;
        }
    }
}

// Decision per candidate (slot k of list S, network outputs y[6 k..]): answered when the most
// probable class is escaped with probability >= pmin and the calibration of the ray's bin allows
// it (held-out rays >= SUR_CAL_MIN, none misclassified, largest direction error <= err_max); the
// answer's direction is the reference's plus the predicted residual. Everything else is appended
// to list A (count counts_a[0]); a seeded subset of answers also to the audit list.
__global__ __launch_bounds__(256) void k_sur_decide(const int32_t* __restrict__ list_s, const uint32_t* __restrict__ count_s,
                                                    const float* __restrict__ y, const float* __restrict__ rec,
                                                    const float* __restrict__ cal_err, const uint32_t* __restrict__ cal_n,
                                                    const uint32_t* __restrict__ cal_bad,
                                                    const float* __restrict__ cal_dmax, float pmin, float err_max,
                                                    double audit_fraction, uint64_t audit_seed, DevOut o, int n_bins,
                                                    int32_t* __restrict__ list_a, uint32_t* __restrict__ count_a,
                                                    int32_t* __restrict__ list_audit, uint32_t* __restrict__ count_audit,
                                                    uint32_t* __restrict__ n_gated) {
    __shared__ uint32_t off[2][4];
    __shared__ uint32_t base[2];
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long below = lane ? ((1ull << lane) - 1ull) : 0ull;
    const uint64_t thr = audit_fraction >= 1.0 ? ~0ull : (uint64_t)(audit_fraction * 18446744073709551615.0);
    const int64_t total = *count_s;
    for (int64_t k0 = (int64_t)blockIdx.x * 256; k0 < total; k0 += (int64_t)gridDim.x * 256) {
        const int64_t k = k0 + threadIdx.x;
        bool to_a = false, audit = false, gated = false;
        int64_t id = -1;
        if (k < total) {
            id = list_s[k];
            const float* yy = y + k * SIRIUS_MLP_OUT;
            const float* rr = rec + id * SUR_REC;
            const float m = fmaxf(yy[0], fmaxf(yy[1], yy[2]));
            const float e0 = expf(yy[0] - m), e1 = expf(yy[1] - m), e2 = expf(yy[2] - m);
            const float p0 = e0 / (e0 + e1 + e2);
            const bool esc = e0 >= e1 && e0 >= e2 && p0 >= pmin;
            const bool cal = sur_bin_usable(SurCal{cal_err, cal_n, cal_bad, cal_dmax, err_max, 0.f, 0.f}, sur_bin_rec(rr),
                                            rr[5]);
            gated = esc && !cal;
            if (esc && cal) {
                uint8_t fl = o.flags[id] | (uint8_t)SIRIUS_FLAG_SURROGATE;
                audit = audit_fraction > 0.0 && sur_hash(audit_seed ^ (uint64_t)id) <= thr;
                if (audit) fl |= (uint8_t)SIRIUS_FLAG_AUDITED;
                o.flags[id] = fl;
                o.cls[id] = (uint8_t)SIRIUS_RAY_ESCAPED;
                float v[3], nn = 0.f;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    v[q] = rr[8 + q] + (float)SIRIUS_MLP_DIR_SCALE * yy[3 + q];
                    nn += v[q] * v[q];
                }
                const float il = 1.f / sqrtf(nn);
                o.theta[id] = acosf(fminf(1.f, fmaxf(-1.f, v[2] * il)));
                o.phi[id] = atan2f(v[1], v[0]);
                o.lam[id] = 0.f;
                o.disk_r[id] = 0.f;
                o.redshift[id] = 0.f;
                o.null_drift[id] = 0.f;
                o.steps[id] = 0u;
                o.rejected[id] = 0u;
                o.landings[id] = 0u;
                for (int q = 0; q < n_bins; ++q) o.intensity[(int64_t)q * o.n_rays + id] = 0.f;
            } else {
                to_a = true;
            }
        }
        const unsigned long long ma = __ballot(to_a), mu = __ballot(audit), mg = __ballot(gated);
        if (lane == 0) {
            off[0][wave] = (uint32_t)__popcll(ma);
            off[1][wave] = (uint32_t)__popcll(mu);
            if (mg) atomicAdd(n_gated, (uint32_t)__popcll(mg));
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            uint32_t s = 0;
            for (int w = 0; w < 4; ++w) {
                const uint32_t t = off[threadIdx.x][w];
                off[threadIdx.x][w] = s;
                s += t;
            }
            base[threadIdx.x] = s ? atomicAdd(threadIdx.x == 0 ? count_a : count_audit, s) : 0u;
        }
        __syncthreads();
        if (to_a) list_a[base[0] + off[0][wave] + __popcll(ma & below)] = (int32_t)id;
        if (audit) list_audit[base[1] + off[1][wave] + __popcll(mu & below)] = (int32_t)id;
        __syncthreads();
    }
}

// Calibration of the error gate on held-out candidates (slot k of `list`, outputs y[6 k..],
// integrator labels in o): for each ray the network would answer (escaped with p >= pmin), its
// bin counts it, records a misclassification, or keeps the largest direction error.
__global__ void k_sur_calibrate(const int32_t* __restrict__ list, int64_t m, const float* __restrict__ y,
                                const float* __restrict__ rec, DevOut o, float pmin, float* __restrict__ cal_err,
                                uint32_t* __restrict__ cal_n, uint32_t* __restrict__ cal_bad,
                                float* __restrict__ cal_dmax) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const int64_t id = list[k];
    const float* yy = y + k * SIRIUS_MLP_OUT;
    const float* rr = rec + id * SUR_REC;
    const float mx = fmaxf(yy[0], fmaxf(yy[1], yy[2]));
    const float e0 = expf(yy[0] - mx), e1 = expf(yy[1] - mx), e2 = expf(yy[2] - mx);
    if (!(e0 >= e1 && e0 >= e2 && e0 / (e0 + e1 + e2) >= pmin)) return;
    const int b = sur_bin_rec(rr);
    if (o.cls[id] != SIRIUS_RAY_ESCAPED) {
        atomicAdd(&cal_bad[b], 1u);
        return;
    }
    double v[3], nn = 0.0;
    for (int q = 0; q < 3; ++q) {
        v[q] = (double)rr[8 + q] + SIRIUS_MLP_DIR_SCALE * (double)yy[3 + q];
        nn += v[q] * v[q];
    }
    const double th = o.theta[id], ph = o.phi[id];
    const double e[3] = {sin(th) * cos(ph), sin(th) * sin(ph), cos(th)};
    double dd = 0.0;
    for (int q = 0; q < 3; ++q) {
        const double t = v[q] / sqrt(nn) - e[q];
        dd += t * t;
    }
    const float err = (float)(2.0 * asin(fmin(1.0, 0.5 * sqrt(dd))));
    atomicAdd(&cal_n[b], 1u);
    atomicMax((unsigned int*)&cal_err[b], __float_as_uint(err));
    atomicMax((unsigned int*)&cal_dmax[b], __float_as_uint(rr[5]));  // distances are positive
}

// Gathers of a candidate subset: network inputs from the records (by ray id) into x[5 k..].
__global__ void k_sur_inputs(const int32_t* __restrict__ list, int64_t m, const float* __restrict__ rec,
                             float* __restrict__ x) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const float* rr = rec + (int64_t)list[k] * SUR_REC;
    for (int q = 0; q < SIRIUS_MLP_IN; ++q) x[k * SIRIUS_MLP_IN + q] = rr[q];
}

// Training batch gather through a permutation of the training list.
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
hipError_t launch_sur_features(const CamConst* cams, int n_cam, const float* tabs, int64_t n, double sur_band,
                               double margin, double r_exit, float* x, float* ref, uint8_t* cand, float* dist,
                               float* rec, hipStream_t st) {
    hipLaunchKernelGGL(k_sur_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, cams, n_cam, tabs, n,
                       (float)sur_band, (float)margin, (float)r_exit, x, ref, cand, dist, rec);
    return hipGetLastError();
}
hipError_t launch_sur_targets(const DevOut& o, const float* ref, int64_t n, int32_t* cls, float* dir, hipStream_t st) {
    hipLaunchKernelGGL(k_sur_targets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, o, ref, n, cls, dir);
    return hipGetLastError();
}
hipError_t launch_sur_gate(const CamConst* cams, int n_cam, const float* tabs, int64_t n, double band, double crit_band,
                           double long_band, double sur_band, double margin, double r_exit, int32_t* list_a,
                           int32_t* list_b, int32_t* list_l, int32_t* list_s, uint32_t* counts, uint8_t* flags,
                           float* rec, const float* cal_err, const uint32_t* cal_n, const uint32_t* cal_bad,
                           const float* cal_dmax, double err_max, float sup_x, float sup_y, hipStream_t st) {
    hipLaunchKernelGGL(k_sur_gate, dim3((unsigned)((n + SG_RAYS - 1) / SG_RAYS)), dim3(256), 0, st, cams, n_cam, tabs, n,
                       (float)band, (float)crit_band, (float)long_band, (float)sur_band, (float)margin, (float)r_exit,
                       list_a, list_b, list_l, list_s, counts, flags, rec,
                       SurCal{cal_err, cal_n, cal_bad, cal_dmax, (float)err_max, sup_x, sup_y});
    return hipGetLastError();
}
hipError_t launch_sur_decide(const int32_t* list_s, const uint32_t* count_s, int64_t cap, const float* y,
                             const float* rec, const float* cal_err, const uint32_t* cal_n, const uint32_t* cal_bad,
                             const float* cal_dmax, double pmin, double err_max, double audit_fraction,
                             uint64_t audit_seed, const DevOut& o,
                             int n_bins, int32_t* list_a, uint32_t* count_a, int32_t* list_audit,
                             uint32_t* count_audit, uint32_t* n_gated, hipStream_t st) {
    int64_t blocks = (cap + 255) / 256;
    blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
    hipLaunchKernelGGL(k_sur_decide, dim3((unsigned)blocks), dim3(256), 0, st, list_s, count_s, y, rec, cal_err, cal_n,
                       cal_bad, cal_dmax, (float)pmin, (float)err_max, audit_fraction, audit_seed, o, n_bins, list_a,
                       count_a, list_audit, count_audit, n_gated);
    return hipGetLastError();
}
hipError_t launch_sur_calibrate(const int32_t* list, int64_t m, const float* y, const float* rec, const DevOut& o,
                                double pmin, float* cal_err, uint32_t* cal_n, uint32_t* cal_bad, float* cal_dmax,
                                hipStream_t st) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_calibrate, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, list, m, y, rec, o,
                       (float)pmin, cal_err, cal_n, cal_bad, cal_dmax);
    return hipGetLastError();
}
hipError_t launch_sur_inputs(const int32_t* list, int64_t m, const float* rec, float* x, hipStream_t st) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sur_inputs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, list, m, rec, x);
    return hipGetLastError();
}
}  // namespace sirius
