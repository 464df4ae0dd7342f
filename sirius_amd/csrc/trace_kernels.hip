// trace_kernels.hip — geodesic tracing kernels for gfx950 (MI355X).
//
// Role replaced: the OptiX raygen mega-kernel's integrate/terminate/shade loop
// (src/Sirius.Render/Acceleration/OptiX/RDOP002A.cu:5132-6106), with the physics contract of
// the CPU tracer (src/Sirius.Render/Integration/GTRC001A.cpp:139-592) instead of the
// OptiX kernel's approximations. Kernels:
//   k_trace_persistent    DP45 tracer (FP64): waves pull batches of rays from an atomic work
//   k_trace_persistent32  counter and refill idle lanes as rays terminate (FP32 variant with its
//                         own register budget). FP32 rays that meet an escalation criterion are
//                         pushed to a queue and re-traced in FP64. The grid schedule is the same
//                         kernel with one wave per 64 rays and no refills.
//   k_finish              per-ray outputs (directions, emission, null drift) in FP64
//   k_symplectic_grid     fixed-step Tao/Yoshida integrator
#include <hip/hip_runtime.h>

#include <algorithm>

#include "geodesic.hpp"
#include "../../include/sirius_hip.h"

namespace sirius {

enum { ST_RUN = 0, ST_DONE = 1, ST_ESCALATE = 2 };

SIRIUS_HD __forceinline__ int find_camera(const CamConst* cams, int n_cam, int64_t id) {
    int lo = 0, hi = n_cam - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cams[mid].offset <= id) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Null drift |g_{mu nu} k^mu k^nu| = |p.k| of the final state, evaluated in FP64.
template <int FAM>
SIRIUS_HD __forceinline__ double null_drift64(const CamConst& c, const double x[4], const double p[4]) {
    Spacetime<double> s = make_spacetime<double>(c);
    double k[4];
    raise_index<double, FAM>(s, x, p, k);
    return fabs(p[0] * k[0] + p[1] * k[1] + p[2] * k[2] + p[3] * k[3]);
}

// FP64 re-imposition of the null constraint on an FP32 state, keeping E = -p_t and the
// direction of the spatial momentum: p_i -> sigma p_i with H(sigma) = 0, i.e.
// A sigma^2 + B sigma + C = 0, A = q.G q, B = 2 p_t (G q)^t, C = p_t^2 G^{tt}.
template <int FAM>
SIRIUS_HD __forceinline__ void renorm_energy64(const Spacetime<double>& s, const float xf[4], float pf[4], float lo[8]) {
    double x[4], q[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < 4; ++i) x[i] = (double)xf[i] + (double)lo[i];
    for (int i = 1; i < 4; ++i) q[i] = (double)pf[i] + (double)lo[4 + i];
    double e[4] = {1.0, 0.0, 0.0, 0.0};
    double kq[4], ke[4];
    raise_index<double, FAM>(s, x, q, kq);
    raise_index<double, FAM>(s, x, e, ke);
    double pt = (double)pf[0] + (double)lo[4];
    double A = q[1] * kq[1] + q[2] * kq[2] + q[3] * kq[3];
    double B = 2.0 * pt * kq[0];
    double C = pt * pt * ke[0];
    double disc = B * B - 4.0 * A * C;
    if (A <= 0.0 || disc < 0.0) return;
    double sq = sqrt(disc);
    double s1 = (-B + sq) / (2.0 * A), s2 = (-B - sq) / (2.0 * A);
    double sg = fabs(s1 - 1.0) < fabs(s2 - 1.0) ? s1 : s2;
    if (!(sg > 0.5 && sg < 2.0)) return;
    for (int i = 1; i < 4; ++i) {
        double v = q[i] * sg;
        pf[i] = (float)v;
        lo[4 + i] = (float)(v - (double)pf[i]);
    }
}

// ------------------------------------------------------------------------------------------
// Per-lane ray state and one step attempt.
// EMIT is the emission model (SIRIUS_EMISSION_*), a template parameter so each kernel only
// carries the code (and registers) of its own model. TIME (FP64 only) integrates the
// coordinate time, which enters the reference's error norm (PHGD001A.cpp:508-517).
// ------------------------------------------------------------------------------------------
// Volumetric emission: up to this many spectral bins accumulate in registers for the ray's whole
// trace (in the path's precision: an FP32 ray's few hundred node contributions add up to well
// inside its 1e-3 bar) and are stored once when it ends; more bins accumulate in device memory.
constexpr int SIRIUS_VOL_REG = 8;

template <class T, int FAM>
struct Ray {
    T x[4], p[4];
    T lo[8];        // FP32: low parts of the compensated state (x^mu, then p_mu); unused in FP64
    Der<T> d;
    T h, lam;
    T tau;          // volumetric optical depth
    T vr_prev;      // radial velocity sign tracking (FP32 escalation)
    T r_cur;        // event radius of the current state (kept from the step that reached it)
    T rdisk;        // disk-hit radius of a finished ray
    int64_t id;
    int cam;
    int cls;        // termination class of a finished ray
    uint32_t attempts, rejected, landings, flags;
    T vacc[SIRIUS_VOL_REG];  // volumetric emission accumulators (n_bins <= SIRIUS_VOL_REG)
};

template <class T, int FAM, int EMIT>
SIRIUS_HD __forceinline__ void ray_start(Ray<T, FAM>& R, const CamConst* cams, int cam, int64_t id, const TraceParams& tp,
                                         const DevOut& o) {
    const CamConst& c = cams[cam];
    int64_t pix = id - c.offset;
    R.id = id;
    R.cam = cam;
    Spacetime<T> s = make_spacetime<T>(c);
    if constexpr (sizeof(T) == 4) {
        // FP32: the initial state in FP64, kept as high + low parts
        Spacetime<double> sd = make_spacetime<double>(c);
        double xd[4], pd[4];
        init_state<double>(c, sd, (int)(pix % c.width), (int)(pix / c.width), xd, pd);
        for (int i = 0; i < 4; ++i) {
            R.x[i] = (T)xd[i];
            R.p[i] = (T)pd[i];
            R.lo[i] = (T)(xd[i] - (double)R.x[i]);
            R.lo[4 + i] = (T)(pd[i] - (double)R.p[i]);
        }
    } else {
        init_state<T>(c, s, (int)(pix % c.width), (int)(pix / c.width), R.x, R.p);
    }
    eval<T, FAM>(s, R.x, R.p, R.d);
    R.r_cur = event_radius(s, R.x);
    R.h = T(tp.h_init);
    if (tp.far_slope > 0) {
        // the first step starts at the far-field cap (with h0 = 0.01 a ray would spend ~10
        // doubling steps reaching it from the observer)
        T cap = R.r_cur * T(tp.far_slope);
        if (cap < T(tp.h_max)) cap = T(tp.h_max);
        if (cap > R.h) R.h = cap;
    }
    R.lam = T(0);
    R.tau = T(0);
    R.vr_prev = T(-1);
    R.attempts = R.rejected = R.landings = 0;
    R.flags = tp.gate_flags ? (o.flags[id] & SIRIUS_FLAG_CRITICAL) : 0;
    if constexpr (EMIT == SIRIUS_EMISSION_VOLUMETRIC) {
#pragma unroll
        for (int k = 0; k < SIRIUS_VOL_REG; ++k) R.vacc[k] = T(0);
        if (tp.n_bins > SIRIUS_VOL_REG)
            for (int k = 0; k < tp.n_bins; ++k) o.vol_acc[(int64_t)k * o.n_rays + id] = 0.0;
    }
}

// A finished ray leaves its class, counters and final state (x^mu, p_mu in FP64) in the
// result arrays; k_finish turns them into directions, emission and null drift afterwards, so
// the tracing kernels carry none of that FP64 code.
template <class T, int FAM, int EMIT>
SIRIUS_HD __forceinline__ void ray_store(const Ray<T, FAM>& R, const DevOut& o, const TraceParams& tp) {
    int64_t id = R.id;
    if constexpr (EMIT == SIRIUS_EMISSION_VOLUMETRIC) {
        if (tp.n_bins <= SIRIUS_VOL_REG) {
#pragma unroll
            for (int k = 0; k < SIRIUS_VOL_REG; ++k)
                if (k < tp.n_bins) o.vol_acc[(int64_t)k * o.n_rays + id] = (double)R.vacc[k];
        }
    }
    o.cls[id] = (uint8_t)R.cls;
    o.flags[id] = (uint8_t)R.flags;
    o.lam[id] = (float)R.lam;
    o.disk_r[id] = (float)R.rdisk;
    o.redshift[id] = EMIT == SIRIUS_EMISSION_VOLUMETRIC ? (float)R.tau : 0.f;
    o.steps[id] = R.attempts;
    o.rejected[id] = R.rejected;
    o.landings[id] = R.landings;
    for (int i = 0; i < 4; ++i) {
        double xi = (double)R.x[i], pi = (double)R.p[i];
        if constexpr (sizeof(T) == 4) { xi += (double)R.lo[i]; pi += (double)R.lo[4 + i]; }
        o.state[8 * id + i] = xi;
        o.state[8 * id + 4 + i] = pi;
    }
}

// Outputs of one finished ray from its stored class and state (all FP64).
template <int FAM, int EMIT>
SIRIUS_HD __forceinline__ void ray_outputs(const CamConst& c, const TraceParams& tp, const DevOut& o, int64_t id) {
    const double* st = o.state + 8 * id;
    double xd[4] = {st[0], st[1], st[2], st[3]}, pd[4] = {st[4], st[5], st[6], st[7]};
    Spacetime<double> sd = make_spacetime<double>(c);
    double kd[4];
    raise_index<double, FAM>(sd, xd, pd, kd);
    int cls = o.cls[id];
    float th = 0.f, ph = 0.f;
    if (cls == SIRIUS_RAY_ESCAPED) {
        double dir[3];
        if constexpr (FAM == 1) {
            // wormhole: map the orbital-plane end state back to the exit side's sky
            int64_t pix = id - c.offset;
            double ns[3];
            camera_dir<double>(c, (int)(pix % c.width), (int)(pix / c.width), ns);
            double nt = sqrt(ns[1] * ns[1] + ns[2] * ns[2]);
            double e2[3];
            for (int i = 0; i < 3; ++i) e2[i] = nt > 1e-300 ? (ns[1] * c.et[i] + ns[2] * c.ep[i]) / nt : c.et[i];
            double l = xd[1], psi = xd[2], r = sqrt(l * l + c.b0 * c.b0);
            double drdl = (l / r) * kd[1], rdpsi = pd[2] / r;
            double cp = cos(psi), sp = sin(psi);
            for (int i = 0; i < 3; ++i) {
                double e = cp * c.er[i] + sp * e2[i], t = -sp * c.er[i] + cp * e2[i];
                dir[i] = drdl * e + rdpsi * t;
            }
            if (l < 0) o.flags[id] |= SIRIUS_FLAG_FAR_SIDE;
        } else {
            dir[0] = kd[1];
            dir[1] = kd[2];
            dir[2] = kd[3];
        }
        double n = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        th = (float)acos(fmin(1.0, fmax(-1.0, dir[2] / n)));
        ph = (float)atan2(dir[1], dir[0]);
    }
    if constexpr (EMIT == SIRIUS_EMISSION_THIN_DISK) {
        if (cls == SIRIUS_RAY_DISK) {
            double gf;
            double rd = kerr_r(sd, xd[1], xd[2], xd[3]);  // FP64 radius of the landed state
            // rays re-traced in FP64 keep FP64 bins
            bool bins32 = tp.spectrum32 && !(o.flags[id] & SIRIUS_FLAG_ESCALATED);
            disk_emission(c, tp, xd, pd, rd, gf, o.intensity + id, (int)o.n_rays, bins32);
            o.redshift[id] = (float)gf;
        } else {
            for (int k = 0; k < tp.n_bins; ++k) o.intensity[(int64_t)k * o.n_rays + id] = 0.f;
        }
    } else if constexpr (EMIT == SIRIUS_EMISSION_VOLUMETRIC) {
        for (int k = 0; k < tp.n_bins; ++k)
            o.intensity[(int64_t)k * o.n_rays + id] = (float)o.vol_acc[(int64_t)k * o.n_rays + id];
    } else {
        for (int k = 0; k < tp.n_bins; ++k) o.intensity[(int64_t)k * o.n_rays + id] = 0.f;
    }
    o.theta[id] = th;
    o.phi[id] = ph;
    o.null_drift[id] = (float)fabs(pd[0] * kd[0] + pd[1] * kd[1] + pd[2] * kd[2] + pd[3] * kd[3]);
}

// Volumetric transfer over an accepted step (3-node Gauss-Legendre on the cubic Hermite
// interpolant of (x, p)). Emitter-frame j_nu = rho S(x_e), alpha = kappa rho, invariant form:
//   dtau = kappa rho nu_e dlam,  dI_k = rho S(x_k / g) (nu_o^3 / nu_e^2) e^{-tau} dlam.
// Smooth (C2) radial window of the emitting shell: zero at r_min and r_max, one between
// 1.1 r_min and 0.9 r_max (quintic smoothstep 6u^5 - 15u^4 + 10u^3), so the emissivity has no
// jump, and no jump in its first two derivatives, for the per-step quadrature to straddle. With
// the cubic (C1) smoothstep the quadrature's error at the shell's edges set how long a step
// could be: on the 64-camera batch, steps of 0.3 sigma left 1.9e-3 of intensity error against
// the FP64 reference with the cubic window, 6.0e-4 with this one.
template <class T>
SIRIUS_HD __forceinline__ T vol_window(T r, T r_min, T r_max) {
    T u = (r - r_min) / (T(0.1) * r_min), v = (r_max - r) / (T(0.1) * r_max);
    u = u < T(0) ? T(0) : (u > T(1) ? T(1) : u);
    v = v < T(0) ? T(0) : (v > T(1) ? T(1) : v);
    return u * u * u * (u * (T(6) * u - T(15)) + T(10)) * v * v * v * (v * (T(6) * v - T(15)) + T(10));
}

template <class T, int FAM>
SIRIUS_HD __forceinline__ void vol_accumulate(Ray<T, FAM>& R, const CamConst& c, const Spacetime<T>& s,
                                              const TraceParams& tp, const DevOut& o, const T x1[4], const T p1[4],
                                              const Der<T>& d1) {
    const T gn[3] = {T(0.1127016653792583), T(0.5), T(0.8872983346207417)};
    const T gw[3] = {T(5.0 / 18.0), T(8.0 / 18.0), T(5.0 / 18.0)};
    double xk0, qk0;
    bin_geometry(tp, xk0, qk0);
    // one node at a time: unrolled, the three nodes' spectra would be live together
#pragma unroll 1
    for (int q = 0; q < 3; ++q) {
        T xq[4];
        hermite(R.x, R.d.k, x1, d1.k, R.h, gn[q], xq);
        T r = kerr_r(s, xq[1], xq[2], xq[3]);
        const T r_lo = T(c.r_in > tp.vol_r_min ? c.r_in : tp.vol_r_min);  // no emitters inside the ISCO
        if (r < r_lo || r > T(tp.vol_r_max)) continue;
        T sig = T(tp.vol_h_over_r) * r;
        // beyond 10 scale heights the density factor is below e^-50: nothing to add
        T zarg = xq[3] * xq[3] / (T(2) * sig * sig);
        if (zarg > T(50)) continue;
        T rho = T(tp.vol_j0) / (r * r) * exp(-zarg) * vol_window(r, r_lo, T(tp.vol_r_max));
        // emitter velocity, frequencies and the spectrum in the path's precision (each node's
        // contribution then joins an FP64 accumulator); p by cubic Hermite with dp/dlam
        T th = gn[q], t2 = th * th, t3 = t2 * th, hh = R.h;
        T h00 = T(2) * t3 - T(3) * t2 + T(1), h10 = t3 - T(2) * t2 + th, h01 = T(3) * t2 - T(2) * t3, h11 = t3 - t2;
        T pq[4];
        pq[0] = R.p[0];
        for (int i = 1; i < 4; ++i)
            pq[i] = h00 * R.p[i] + h10 * hh * R.d.f[i - 1] + h01 * p1[i] + h11 * hh * d1.f[i - 1];
        T rc = sqrt(fmax(xq[1] * xq[1] + xq[2] * xq[2] - T(c.a * c.a), T(1e-12)));
        T Om = kepler_omega<T>(c, rc);
        T ut = emitter_ut<T>(c, s, xq, Om);
        if (ut < T(0)) continue;
        T nue = -ut * (pq[0] + Om * (xq[1] * pq[2] - xq[2] * pq[1]));
        T nuo = -pq[0] * T(c.obs_norm);
        if (nue <= T(0) || nuo <= T(0)) continue;
        T gf = nuo / nue;
        T dl = gw[q] * hh;
        T dtau = T(tp.vol_kappa) * rho * nue * dl;
        T w = rho * (nuo * nuo * nuo / (nue * nue)) * dl * exp(-(R.tau + T(0.5) * dtau));
        T tv = pow_pos(T(6) / r, T(0.75));
        T xk = T(xk0), qk = T(qk0);
        if (tp.n_bins <= SIRIUS_VOL_REG) {
#pragma unroll
            for (int k = 0; k < SIRIUS_VOL_REG; ++k) {
                if (k < tp.n_bins) {
                    T xe = xk / gf;
                    R.vacc[k] += w * xe * xe * xe / expm1(xe / tv);
                }
                xk *= qk;
            }
        } else {
            for (int k = 0; k < tp.n_bins; ++k, xk *= qk) {
                T xe = xk / gf;
                o.vol_acc[(int64_t)k * o.n_rays + R.id] += (double)(w * xe * xe * xe / expm1(xe / tv));
            }
        }
        R.tau += dtau;
    }
}

template <class T>
SIRIUS_HD __forceinline__ T vol_step_cap(const Spacetime<T>& s, const TraceParams& tp, const T x[4]) {
    T r = kerr_r(s, x[1], x[2], x[3]);
    if (r < T(0.5 * tp.vol_r_min) || r > T(1.5 * tp.vol_r_max)) return T(1e30);
    T sig = T(tp.vol_h_over_r) * r;
    T cap = T(0.5) * (fabs(x[3]) - T(3) * sig);
    // inside the layer at most 0.2 sigma per step (FP64, the oracle's), or 0.3 sigma on the FP32
    // path, whose intensity tolerance is 1e-3: 0.3 sigma keeps the 64-camera batch within
    // 6.0e-4 of the FP64 reference and takes 95 instead of 118 steps per ray
    const T inner = T(sizeof(T) == 4 ? 0.3 : 0.2);
    return cap > inner * sig ? cap : inner * sig;
}

// Land exactly on an event at fraction th of the step: one DP45 step of th*h from the step's
// start, and for the disk plane one Newton correction of th. Landings always end the ray, so
// the state is overwritten with the landed one. Returns the final th.
template <class T, int FAM>
SIRIUS_HD __forceinline__ T land(Ray<T, FAM>& R, const Spacetime<T>& s, bool plane, T th) {
    constexpr bool TIME = sizeof(T) == 8;
    T xl[4], pl[4], err[4];
    Der<T> dl;
    dp45_step<T, FAM, TIME>(s, R.x, R.p, R.d, th * R.h, xl, pl, dl, err, false);
    R.landings++;
    if (plane && fabs(dl.k[3]) > T(1e-30)) {
        T th2 = th - xl[3] / (dl.k[3] * R.h);
        if (th2 > T(0) && th2 < T(1.5)) {
            th = th2;
            dp45_step<T, FAM, TIME>(s, R.x, R.p, R.d, th * R.h, xl, pl, dl, err, false);
            R.landings++;
        }
    }
    R.lam += th * R.h;
    for (int i = 0; i < 4; ++i) { R.x[i] = xl[i]; R.p[i] = pl[i]; }
    R.d = dl;
    return th;
}

// The step's low parts become the state's (t and p_t are not stepped: their low parts stay).
template <class T>
SIRIUS_HD __forceinline__ void commit_lo(T lo[8], const T lo1[8]) {
#pragma unroll
    for (int j = 1; j < 8; ++j)
        if (j != 4) lo[j] = lo1[j];
}

// FP32 landing: the state at fraction th of the step from the step's dense output (x, y, z and
// p_x, p_y, p_z; t and p_t are not stepped in FP32). Landings end the ray.
template <class T, int FAM>
SIRIUS_HD __forceinline__ void land_dense(Ray<T, FAM>& R, const T x1[4], const T p1[4], const Der<T>& d1,
                                          const T dn[6], const T lo1[8], T th) {
    T xl[4], pl[4];
    dense_pos(R.x, R.d.k, x1, d1.k, dn, R.h, th, xl);
    pl[0] = R.p[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) pl[i] = dense_eval(R.p[i], p1[i], R.h, R.d.f[i - 1], d1.f[i - 1], dn[2 + i], th);
    R.lam = tfma(th, R.h, R.lam);
    for (int i = 0; i < 4; ++i) { R.x[i] = xl[i]; R.p[i] = pl[i]; }
    // low parts of the compensated state, linear across the step (they are below the high
    // parts' rounding unit)
#pragma unroll
    for (int j = 1; j < 8; ++j)
        if (j != 4) R.lo[j] = tfma(th, lo1[j] - R.lo[j], R.lo[j]);
}

// One DP45 attempt with error control, events and acceptance. Returns ST_RUN, ST_DONE (the
// ray ended: R.cls, R.rdisk and the final state are set; the caller writes it out with
// ray_finish) or ST_ESCALATE (an FP32 lane hands the ray to the FP64 queue). Every event is
// located and landed through one code path, so the landing code exists once per kernel.
template <class T, int FAM, int EMIT, bool MULTI>
SIRIUS_HD __forceinline__ int ray_attempt(Ray<T, FAM>& R, const CamConst* cams, const Spacetime<T>& s,
                                          const TraceParams& tp, const DevOut& o) {
    constexpr bool fp32 = sizeof(T) == 4;
    constexpr bool TIME = !fp32;
    constexpr bool vol = (FAM == 0) && EMIT == SIRIUS_EMISSION_VOLUMETRIC;
    T hmax = T(tp.h_max);
    T r0 = T(0);
    if (tp.far_slope > 0 || tp.step_rexp > 0) r0 = R.r_cur;
    if (tp.far_slope > 0) {
        T rr = r0 * T(tp.far_slope);
        hmax = rr > hmax ? rr : hmax;
    }
    if constexpr (vol) {
        T cap = vol_step_cap(s, tp, R.x);
        if (R.h > cap) R.h = cap;
    }
    T x1[4], p1[4], err[4];
    Der<T> d1;
    // FP32: the step's dense output locates and lands events (no further evaluations)
    T dn[6], lo1[8];
    dp45_step<T, FAM, TIME>(s, R.x, R.p, R.d, R.h, x1, p1, d1, err, true, fp32 ? dn : nullptr,
                            fp32 ? R.lo : nullptr, lo1);
    R.attempts++;
    T e = T(0);
#pragma unroll
    for (int i = TIME ? 0 : 1; i < 4; ++i) {
        T q = err[i] / (T(tp.atol) + T(tp.rtol) * fabs(x1[i]));
        e += q * q;
    }
    e = sqrt(e * (TIME ? T(0.25) : T(1.0 / 3.0)));
    if (!(e <= T(1))) {
        R.h = next_step(R.h, e, T(tp.h_min), hmax);
        R.rejected++;
        if (!sfinite(e) || R.h <= T(tp.h_min)) { R.cls = SIRIUS_RAY_FAILED; return ST_DONE; }
        if (R.attempts >= (uint32_t)tp.max_steps) { R.cls = SIRIUS_RAY_MAX_STEPS; return ST_DONE; }
        return ST_RUN;
    }
    bool bad = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) bad |= !sfinite(x1[i]) || !sfinite(p1[i]);
    if (bad) { R.cls = SIRIUS_RAY_FAILED; return ST_DONE; }
    const CamConst& c = cams[MULTI ? R.cam : 0];
    // ---- volumetric transfer ----
    int ev = -1;        // pending termination class
    int kind = EV_R;    // event function to locate (EV_Z plane or EV_R radius)
    T target = T(0);
    bool locate = false;
    if constexpr (vol) {
        vol_accumulate<T, FAM>(R, c, s, tp, o, x1, p1, d1);
        if (R.tau > T(tp.vol_tau_max)) ev = SIRIUS_RAY_DISK;
    }
    // ---- thin disk: first crossing of the equatorial plane inside [r_in, r_out], decided on
    // the Hermite-interpolated crossing radius, reported from the landed state ----
    T th_disk = T(1);
    if constexpr (FAM == 0 && EMIT == SIRIUS_EMISSION_THIN_DISK) {
        if (R.x[3] * x1[3] < T(0)) {
            T xc[4];
            if constexpr (fp32) {
                th_disk = locate_event_newton(s, EV_Z, T(0), R.x, R.d.k, x1, d1.k, dn, R.h, R.x[3], x1[3]);
                dense_pos(R.x, R.d.k, x1, d1.k, dn, R.h, th_disk, xc);
            } else {
                th_disk = locate_event(s, EV_Z, T(0), R.x, R.d.k, x1, d1.k, R.h);
                hermite(R.x, R.d.k, x1, d1.k, R.h, th_disk, xc);
            }
            T rc = kerr_r(s, xc[1], xc[2], xc[3]);
            if constexpr (fp32) {
                // near the inner edge the emission falls steeply to zero: an FP32 position error
                // there is a large relative intensity error, so such rays go to FP64 (rays in the
                // critical band, which the accuracy contract exempts, stay FP32)
                if (tp.escalate && !(R.flags & SIRIUS_FLAG_CRITICAL) && fabs(rc - T(c.r_in)) < T(tp.esc_edge * c.r_in))
                    return ST_ESCALATE;
            }
            if (rc >= T(c.r_in) && rc <= T(c.r_out)) ev = SIRIUS_RAY_DISK;
        }
    }
    T r1 = event_radius(s, x1);
    if (ev < 0) {
        if (FAM == 0 && r1 < T(c.r_cap)) {
            ev = SIRIUS_RAY_CAPTURED;
            target = T(c.r_cap);
            locate = event_radius(s, R.x) > target;
        } else if (r1 > T(tp.r_escape)) {
            T radial = (FAM == 1) ? x1[1] * d1.k[1] : x1[1] * d1.k[1] + x1[2] * d1.k[2] + x1[3] * d1.k[3];
            if (radial > T(0)) {
                ev = SIRIUS_RAY_ESCAPED;
                target = T(tp.r_escape);
                locate = event_radius(s, R.x) < target;
            }
        }
    }
    if (ev >= 0) {
        bool plane = (ev == SIRIUS_RAY_DISK) && EMIT == SIRIUS_EMISSION_THIN_DISK;
        if (plane || locate) {
            if constexpr (fp32) {
                T th = plane ? th_disk
                             : locate_event_newton(s, kind, target, R.x, R.d.k, x1, d1.k, dn, R.h, R.r_cur - target,
                                                   r1 - target);
                land_dense<T, FAM>(R, x1, p1, d1, dn, lo1, th);
            } else {
                T th = plane ? th_disk : locate_event(s, kind, target, R.x, R.d.k, x1, d1.k, R.h);
                land<T, FAM>(R, s, plane, th);
            }
        } else {
            R.lam += R.h;
            for (int i = 0; i < 4; ++i) { R.x[i] = x1[i]; R.p[i] = p1[i]; }
            if constexpr (fp32) commit_lo(R.lo, lo1);
            R.d = d1;
        }
        R.cls = ev;
        R.rdisk = plane ? kerr_r(s, R.x[1], R.x[2], R.x[3]) : T(0);
        return ST_DONE;
    }
    // ---- accept ----
    R.lam += R.h;
    T hmax2 = T(tp.h_max);
    if (tp.far_slope > 0) {
        T rr = r1 * T(tp.far_slope);
        hmax2 = rr > hmax2 ? rr : hmax2;
    }
    T hprop = R.h;
    // an inward step predicts the next step's error growth: the local error of a step of a
    // given length rises steeply as r falls, so the proposal scales with (r1 / r0)^exp
    if (tp.step_rexp > 0 && r1 < r0) hprop *= pow_pos(r1 / r0, T(tp.step_rexp));
    R.h = next_step(hprop, e, T(tp.h_min), hmax2);
    R.r_cur = r1;
    for (int i = 0; i < 4; ++i) { R.x[i] = x1[i]; R.p[i] = p1[i]; }
    if constexpr (fp32) commit_lo(R.lo, lo1);
    R.d = d1;
    if constexpr (fp32) {
        // escalation of a warp ray entering the bubble wall (|r_s - R| < 3 / sigma, or inside):
        // rays crossing the wall are deflected by up to ~2 rad and their outcome is sensitive
        // far beyond what the FP32 path's tolerance controls
        if (FAM == 2 && tp.escalate && r1 < T(c.wR + 3.0 / c.wsig)) return ST_ESCALATE;
        // escalation of a wormhole ray whose impact parameter L = p_psi lies within 10 % of the
        // throat radius: it winds about the throat's unstable photon orbit, where FP32 round-off
        // grows exponentially with each turn
        if (FAM == 1 && tp.escalate && fabs(R.p[2] - T(c.b0)) < T(0.1 * c.b0)) return ST_ESCALATE;
        // escalation: a radial turning point inside the photon region
        if (FAM == 0 && tp.escalate) {
            T vr = x1[1] * d1.k[1] + x1[2] * d1.k[2] + x1[3] * d1.k[3];
            if (R.vr_prev < T(0) && vr >= T(0) && r1 < T(tp.escalate_margin * c.r_ph_retro)) return ST_ESCALATE;
            R.vr_prev = vr;
        }
        // FP64 re-imposition of the null constraint, energy preserving, once the drift exceeds
        // 1e-5 p_t^2 (FP32 rounding alone drifts ~1e-7 per step; at the FP64 path's 1e-6 a fifth
        // of the rays re-normalised, and a wave runs the branch whenever one of its lanes does,
        // while at 1e-5 0.6 % do, with the same results to within the accuracy bars)
        if (fabs(R.d.pk) > T(1e-5) * R.p[0] * R.p[0]) {
            Spacetime<double> sd = make_spacetime<double>(c);
            renorm_energy64<FAM>(sd, R.x, R.p, R.lo);
            eval<T, FAM>(s, R.x, R.p, R.d);
            R.flags |= SIRIUS_FLAG_RENORM;
        }
    } else if (fabs(R.d.pk) > T(1e-6)) {
        // FP64 path: reference re-normalisation, k^0 re-solved with k^i fixed (PHGD001A.cpp:609-638),
        // taking the root nearest the current k^0
        T g[4][4], k[4] = {R.d.k[0], R.d.k[1], R.d.k[2], R.d.k[3]};
        metric_cov(s, R.x, g);
        T k0 = null_k0(g, k, k[0]);
        if (sfinite(k0)) {
            k[0] = k0;
            for (int mu = 0; mu < 4; ++mu) {
                T acc = T(0);
                for (int nu = 0; nu < 4; ++nu) acc += g[mu][nu] * k[nu];
                R.p[mu] = acc;
            }
            eval<T, FAM>(s, R.x, R.p, R.d);
            R.flags |= SIRIUS_FLAG_RENORM;
        }
    }
    if (R.attempts >= (uint32_t)tp.max_steps) { R.cls = SIRIUS_RAY_MAX_STEPS; return ST_DONE; }
    return ST_RUN;
}

// ------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------

// Slots of the work list are reserved per wave from the queue head. One atomic on a single
// address per refill stalls the refilling wave for the atomic's round trip, which under the
// contention of a whole-chip grid costs more than the refill itself (1080p Kerr frame, FP32:
// 1.33 ms at a refill threshold of 32, 2.56 ms at 2). A wave therefore reserves SLOT_CHUNK
// slots at a time and serves its refills from them, until the head comes within SLOT_TAIL_WAVES
// chunks per wave of the list's end; from there on it reserves exactly its idle lanes, so no
// wave sits on unstarted rays while others have drained.
constexpr int64_t SLOT_CHUNK = 128;
constexpr int64_t SLOT_TAIL_WAVES = 2;
static_assert(SLOT_CHUNK >= 64, "a reservation covers a whole wave's refill");

// Persistent tracer. Each wave serves its refills from reserved work-list slots (above) and
// refills its idle lanes when at least refill_min of them are idle; slot -> ray through ray_list
// when given. FP32 lanes that must escalate append their ray to esc_list (count in *esc_count);
// each lane adds the attempts of every ray it traced to *attempts_out when the wave retires.
template <class T, int FAM, int EMIT, bool MULTI>
__device__ __forceinline__ void persistent_body(const CamConst* __restrict__ cams, TraceParams tp, DevOut o,
                                                          const int32_t* __restrict__ ray_list,
                                                          const uint32_t* __restrict__ list_count, int64_t n,
                                                          uint32_t* __restrict__ head, uint32_t* __restrict__ esc_count,
                                                          int32_t* __restrict__ esc_list,
                                                          unsigned long long* __restrict__ attempts_out,
                                                          int refill_min, uint32_t flag_or) {
    const int lane = __lane_id();
    int64_t main_total = list_count ? (int64_t)*list_count : n;
    if (main_total > n) main_total = n;
    int64_t first_total = tp.first_count ? (int64_t)*tp.first_count : 0;
    if (first_total > n) first_total = n;
    const int64_t total = main_total + first_total;
    Ray<T, FAM> R;
    Spacetime<T> s;
    if (!MULTI) s = make_spacetime<T>(cams[0]);
    bool active = false;
    bool exhausted = false;    // no slot left for this wave: neither reserved nor in the queue
    bool queue_done = false;   // the queue head has passed the list's end
    // slots are 32-bit (work lists hold int32 ray ids); the wave's reserved slots [res_lo, res_hi)
    uint32_t res_lo = 0, res_hi = 0, last_base = 0;
    const int64_t tail_from = total - SLOT_TAIL_WAVES * SLOT_CHUNK * ((int64_t)gridDim.x * blockDim.x / 64);
    const uint32_t tail_start = tail_from > 0 ? (uint32_t)tail_from : 0u;
    const uint32_t total32 = (uint32_t)total;
    unsigned long long my_attempts = 0;  // attempts of every ray this lane traced
    while (true) {
        unsigned long long idle = __ballot(!active);
        int n_idle = __popcll(idle);
        if (n_idle == 64 && exhausted) break;
        if (!exhausted && (n_idle >= refill_min || n_idle == 64)) {
            // idle lane of rank r takes reserved slot res_lo + r while reserved ones last, then
            // slot base + (r - from_old) of a new reservation
            const uint32_t old_lo = res_lo;
            const uint32_t ni = (uint32_t)n_idle;
            const uint32_t from_old = res_hi - res_lo < ni ? res_hi - res_lo : ni;
            const bool reserve = from_old < ni && !queue_done;
            uint32_t base = 0;
            if (reserve) {
                const uint32_t want = last_base < tail_start ? (uint32_t)SLOT_CHUNK : ni;
                // every lane of the wave is here (the branch is wave-uniform), so lane 0 reserves
                // and the base is read back as a scalar: the reservation state stays in SGPRs
                uint32_t b = 0;
                if (lane == 0) b = atomicAdd(head, want);
                base = __builtin_amdgcn_readfirstlane(b);
                last_base = base;
                if ((uint64_t)base + want >= (uint64_t)total32) queue_done = true;
                res_lo = base + (ni - from_old);
                res_hi = base + want;
            } else {
                res_lo += from_old;
            }
            if (queue_done && (res_hi < total32 ? res_hi : total32) <= res_lo) exhausted = true;
            if (!active) {
                unsigned long long below = lane ? (idle & ((1ull << lane) - 1ull)) : 0ull;
                const uint32_t r = (uint32_t)__popcll(below);
                const int64_t slot = r < from_old ? (int64_t)old_lo + r
                                                  : (reserve ? (int64_t)base + (r - from_old) : total);
                if (slot < total) {
                    int64_t id = slot < first_total ? (int64_t)tp.first_list[slot]
                                                    : (ray_list ? (int64_t)ray_list[slot - first_total] : slot - first_total);
                    int cam = MULTI ? find_camera(cams, tp.n_cam, id) : 0;
                    ray_start<T, FAM, EMIT>(R, cams, cam, id, tp, o);
                    R.flags |= flag_or;
                    if (MULTI) s = make_spacetime<T>(cams[cam]);
                    active = true;
                }
            }
        }
        if (active) {
            int st = ray_attempt<T, FAM, EMIT, MULTI>(R, cams, s, tp, o);
            if (st == ST_ESCALATE) {
                uint32_t k = atomicAdd(esc_count, 1u);
                esc_list[k] = (int32_t)R.id;
            } else if (st == ST_DONE) {
                ray_store<T, FAM, EMIT>(R, o, tp);
            }
            if (st != ST_RUN) {
                my_attempts += R.attempts;
                active = false;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) my_attempts += __shfl_xor(my_attempts, off);
    if (lane == 0) atomicAdd(attempts_out, my_attempts);
}


#define SIRIUS_PERSISTENT_PARAMS                                                                           \
    const CamConst *__restrict__ cams, TraceParams tp, DevOut o, const int32_t *__restrict__ ray_list,     \
        const uint32_t *__restrict__ list_count, int64_t n, uint32_t *__restrict__ head,                   \
        uint32_t *__restrict__ esc_count, int32_t *__restrict__ esc_list,                                  \
        unsigned long long *__restrict__ attempts_out, int refill_min, uint32_t flag_or
#define SIRIUS_PERSISTENT_ARGS \
    cams, tp, o, ray_list, list_count, n, head, esc_count, esc_list, attempts_out, refill_min, flag_or

// FP64 variants: register allocation left to the compiler (256 VGPRs plus 6-156 AGPRs, one wave
// per SIMD). Held to the 256 registers of two waves, the thin-disk tracer spills 34 to scratch
// and the 1920x1080 Kerr frame slows from 11.3 to 14.9 ms. The single-camera warp tracer, which
// re-traces wall crossers (a tenth of a warp frame, too few rays to hide one wave's latency), is
// held to two waves: 268 -> 256 registers with 6 spilled, and the 1024x1024 warp frame goes from
// 2.79 to 2.54 ms.
constexpr int fp64_min_waves(int fam, bool multi) { return fam == 2 && !multi ? 2 : 1; }
// The persistent tracers start on a 4 KiB boundary of the code object. Their loops are about
// 19 KiB of instructions, and where the linker happened to place them changed their speed: the
// FP32 thin-disk tracer, byte-identical in two builds, ran 1.06-1.16 ms per 1080p launch at one
// address and 1.25-1.59 ms at another (an unrelated edit elsewhere in the file moved it), and
// page-aligned it ran at the fast figure in both builds.
template <class T, int FAM, int EMIT, bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fp64_min_waves(FAM, MULTI)), aligned(4096)))
void k_trace_persistent(SIRIUS_PERSISTENT_PARAMS) {
    persistent_body<T, FAM, EMIT, MULTI>(SIRIUS_PERSISTENT_ARGS);
}

// FP32 variants. The thin-disk tracers, whose natural allocation lies just above the 168 VGPRs
// of three waves per SIMD, are held to that budget (no spills at 168; three waves instead of two).
// The single-camera volumetric tracer's natural allocation (264 registers) leaves one wave per
// SIMD; held to the 256 of two waves it spills 5 and runs a 1024x1024 frame 1.54 times faster
// (30.0 -> 19.5 ms, same results). The others keep the compiler's allocation (the multi-camera
// volumetric tracer takes 256 by itself).
constexpr int fp32_min_waves(int emit, bool multi) {
    return emit == SIRIUS_EMISSION_THIN_DISK ? 3 : (emit == SIRIUS_EMISSION_VOLUMETRIC && !multi ? 2 : 1);
}
template <int FAM, int EMIT, bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fp32_min_waves(EMIT, MULTI)), aligned(4096)))
void k_trace_persistent32(SIRIUS_PERSISTENT_PARAMS) {
    persistent_body<float, FAM, EMIT, MULTI>(SIRIUS_PERSISTENT_ARGS);
}

// Outputs of every traced ray (rays the surrogate answered are skipped).
template <int FAM, int EMIT>
__global__ __launch_bounds__(256) void k_finish(const CamConst* __restrict__ cams, TraceParams tp, DevOut o) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= o.n_rays) return;
    if (o.flags[id] & SIRIUS_FLAG_SURROGATE) return;
    int cam = tp.n_cam == 1 ? 0 : find_camera(cams, tp.n_cam, id);
    ray_outputs<FAM, EMIT>(cams[cam], tp, o, id);
}

// Angle between two directions given as (theta, phi), in FP64 by the chord (accurate down to
// 1e-15 rad, where the arc cosine of a dot product loses everything below ~1e-8).
__device__ __forceinline__ double dir_angle(float ta, float pa, float tb, float pb) {
    const double a[3] = {sin((double)ta) * cos((double)pa), sin((double)ta) * sin((double)pa), cos((double)ta)};
    const double b[3] = {sin((double)tb) * cos((double)pb), sin((double)tb) * sin((double)pb), cos((double)tb)};
    const double d = sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
    return 2.0 * asin(fmin(1.0, 0.5 * d));
}

// Surrogate audit: outputs of the FP64 re-traces of audited rays (in oa), compared with the
// surrogate's answers in o. acc[0]: largest escape-direction error (float bits), acc[1]:
// class mismatches.
__global__ void k_sur_audit(const CamConst* __restrict__ cams, TraceParams tp, DevOut oa, DevOut o,
                            const int32_t* __restrict__ list, const uint32_t* __restrict__ count, int64_t cap,
                            unsigned int* __restrict__ acc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t lim = min((int64_t)*count, cap);
    if (i >= lim) return;
    const int64_t id = list[i];
    const int cam = tp.n_cam == 1 ? 0 : find_camera(cams, tp.n_cam, id);
    ray_outputs<0, SIRIUS_EMISSION_THIN_DISK>(cams[cam], tp, oa, id);
    const int ca = oa.cls[id], cs = o.cls[id];
    if (ca != cs) {
        atomicAdd(&acc[1], 1u);
        return;
    }
    if (ca == SIRIUS_RAY_ESCAPED) {
        const float err = (float)dir_angle(oa.theta[id], oa.phi[id], o.theta[id], o.phi[id]);
        atomicMax(&acc[0], __float_as_uint(err));
    }
}

// Selection of the precision audit: rays of the batch with hash(seed ^ id) below the fraction's
// threshold that the FP32 path traced (not answered by the surrogate) outside the critical band
// (the rays the accuracy contract covers), into list (count *count). Each wave takes a span of
// 64 x PREC_SELECT_SPAN consecutive rays: it counts its picks, reserves their places with one
// atomic (a single-address atomic per wave of 64 rays cost more than the hashing), then writes
// them in ray order.
constexpr int PREC_SELECT_SPAN = 32;
__device__ __forceinline__ bool precision_pick(const DevOut& o, int64_t n, uint64_t thr, uint64_t seed, int64_t id) {
    if (id >= n) return false;
    uint64_t z = (seed ^ (uint64_t)id) + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return z <= thr && !(o.flags[id] & (SIRIUS_FLAG_SURROGATE | SIRIUS_FLAG_CRITICAL));
}
__global__ __launch_bounds__(256) void k_precision_select(const DevOut o, int64_t n, uint64_t thr, uint64_t seed,
                                                          int32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const int lane = __lane_id();
    const int64_t first = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * PREC_SELECT_SPAN;
    if (first >= n) return;  // whole waves
    uint32_t total = 0;
    for (int k = 0; k < PREC_SELECT_SPAN; ++k)
        total += (uint32_t)__popcll(__ballot(precision_pick(o, n, thr, seed, first + 64 * k + lane)));
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(count, total);
    base = __shfl(base, 0);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int k = 0; k < PREC_SELECT_SPAN; ++k) {
        const int64_t id = first + 64 * k + lane;
        const bool take = precision_pick(o, n, thr, seed, id);
        const unsigned long long m = __ballot(take);
        if (take) list[base + __popcll(m & below)] = (int32_t)id;
        base += (uint32_t)__popcll(m);
    }
}

// Largest value over the wave (non-negative floats compare as their bit patterns).
__device__ __forceinline__ unsigned int wave_max_u32(unsigned int v) {
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned int w = (unsigned int)__shfl_xor((int)v, off);
        v = w > v ? w : v;
    }
    return v;
}

// Comparison of the precision audit: the FP64 re-traces (oa) against the batch's results (o).
// acc[0] class mismatches, acc[1] audited rays, acc[2] largest direction error (float bits),
// acc[3] largest relative intensity error (float bits); hist[0..63] and hist[64..127] bin the
// direction and intensity errors by log10, 1/4 decade per bin from 1e-12, for their medians.
// A small grid strides over the list (its length is known only on the device). Each block
// bins in LDS and adds its non-zero bins to hist; counts and maxima are reduced over each wave:
// the errors crowd into a few bins, where one global atomic per ray serialised.
constexpr int PREC_COMPARE_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_precision_compare(const CamConst* __restrict__ cams, TraceParams tp, DevOut oa,
                                                           DevOut o, const int32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ count, int64_t cap,
                                                           unsigned int* __restrict__ acc, unsigned int* __restrict__ hist) {
    __shared__ unsigned int sh[128];
    for (int k = threadIdx.x; k < 128; k += blockDim.x) sh[k] = 0u;
    __syncthreads();
    const int64_t lim = min((int64_t)*count, cap);
    unsigned int n_valid = 0u, n_mis = 0u, dir_bits = 0u, int_bits = 0u;
    auto bin = [](double e) {
        int b = e > 0.0 ? (int)((log10(e) + 12.0) * 4.0) : 0;
        return b < 0 ? 0 : (b > 63 ? 63 : b);
    };
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < lim; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t id = list[i];
        const int cam = tp.n_cam == 1 ? 0 : find_camera(cams, tp.n_cam, id);
        if (cams[cam].family == 1) ray_outputs<1, SIRIUS_EMISSION_NONE>(cams[cam], tp, oa, id);
        else if (cams[cam].family == 2) ray_outputs<2, SIRIUS_EMISSION_NONE>(cams[cam], tp, oa, id);
        else if (tp.emission == SIRIUS_EMISSION_THIN_DISK) ray_outputs<0, SIRIUS_EMISSION_THIN_DISK>(cams[cam], tp, oa, id);
        else ray_outputs<0, SIRIUS_EMISSION_NONE>(cams[cam], tp, oa, id);
        ++n_valid;
        const int ca = oa.cls[id], cs = o.cls[id];
        if (ca != cs) {
            ++n_mis;
        } else if (ca == SIRIUS_RAY_ESCAPED) {
            const double err = dir_angle(oa.theta[id], oa.phi[id], o.theta[id], o.phi[id]);
            const unsigned int eb = __float_as_uint((float)err);
            dir_bits = eb > dir_bits ? eb : dir_bits;
            atomicAdd(&sh[bin(err)], 1u);
        } else if (ca == SIRIUS_RAY_DISK && tp.emission == SIRIUS_EMISSION_THIN_DISK && tp.n_bins > 0) {
            double imax = 0.0, dmax = 0.0;
            for (int b = 0; b < tp.n_bins; ++b) {
                const double ia = oa.intensity[(int64_t)b * oa.n_rays + id], is = o.intensity[(int64_t)b * o.n_rays + id];
                imax = fmax(imax, ia);
                dmax = fmax(dmax, fabs(is - ia));
            }
            if (imax > 0.0) {
                const double rel = dmax / imax;
                const unsigned int rb = __float_as_uint((float)rel);
                int_bits = rb > int_bits ? rb : int_bits;
                atomicAdd(&sh[64 + bin(rel)], 1u);
            }
        }
    }
    // every lane of the block is here: reduce over the wave, then flush the block's bins
    for (int off = 32; off > 0; off >>= 1) {
        n_valid += (unsigned int)__shfl_xor((int)n_valid, off);
        n_mis += (unsigned int)__shfl_xor((int)n_mis, off);
    }
    dir_bits = wave_max_u32(dir_bits);
    int_bits = wave_max_u32(int_bits);
    if (__lane_id() == 0) {
        if (n_valid) atomicAdd(&acc[1], n_valid);
        if (n_mis) atomicAdd(&acc[0], n_mis);
        if (dir_bits) atomicMax(&acc[2], dir_bits);
        if (int_bits) atomicMax(&acc[3], int_bits);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 128; k += blockDim.x)
        if (sh[k]) atomicAdd(&hist[k], sh[k]);
}

// Fill a small device array (graph-safe replacement for a memset).
__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------------
// Symplectic integrator: Tao (2016) extended phase space (q, p, x, y), Strang split
// A(h/2) B(h/2) C(h) B(h/2) A(h/2), composed to 4th order with the Yoshida triple jump
// (weights PHSI001A.h:43-47). Fixed step; disk crossing by linear interpolation.
// ------------------------------------------------------------------------------------------
template <class T, int FAM>
SIRIUS_HD __forceinline__ void tao_a(const Spacetime<T>& s, T d, T q[4], T p[4], T x[4], const T y[4]) {
    Der<T> D;
    eval<T, FAM>(s, q, y, D);
    for (int i = 0; i < 3; ++i) p[i + 1] += d * D.f[i];
    for (int i = 0; i < 4; ++i) x[i] += d * D.k[i];
}
template <class T, int FAM>
SIRIUS_HD __forceinline__ void tao_b(const Spacetime<T>& s, T d, T q[4], const T p[4], const T x[4], T y[4]) {
    Der<T> D;
    eval<T, FAM>(s, x, p, D);
    for (int i = 0; i < 4; ++i) q[i] += d * D.k[i];
    for (int i = 0; i < 3; ++i) y[i + 1] += d * D.f[i];
}
template <class T>
SIRIUS_HD __forceinline__ void tao_c(T cw, T sw, T q[4], T p[4], T x[4], T y[4]) {
    for (int i = 0; i < 4; ++i) {
        T sq = q[i] + x[i], sp = p[i] + y[i], dq = q[i] - x[i], dp = p[i] - y[i];
        T ndq = cw * dq + sw * dp, ndp = -sw * dq + cw * dp;
        q[i] = T(0.5) * (sq + ndq); p[i] = T(0.5) * (sp + ndp);
        x[i] = T(0.5) * (sq - ndq); y[i] = T(0.5) * (sp - ndp);
    }
}

template <class T, int FAM>
__global__ __launch_bounds__(256) void k_symplectic_grid(const CamConst* __restrict__ cams, TraceParams tp, DevOut o) {
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= tp.n_rays) return;
    int cam = tp.n_cam == 1 ? 0 : find_camera(cams, tp.n_cam, id);
    Ray<T, FAM> R;
    ray_start<T, FAM, SIRIUS_EMISSION_THIN_DISK>(R, cams, cam, id, tp, o);
    const CamConst& c = cams[cam];
    Spacetime<T> s = make_spacetime<T>(c);
    T q[4], p[4], x[4], y[4];
    for (int i = 0; i < 4; ++i) { q[i] = x[i] = R.x[i]; p[i] = y[i] = R.p[i]; }
    const T w = T(20);
    const T h = T(tp.sym_h);
    const T wts[3] = {T(1.3512071919596576), T(-1.7024143839193153), T(1.3512071919596576)};
    T cw[3], sw[3];
    for (int j = 0; j < 3; ++j) { cw[j] = cos(T(2) * w * wts[j] * h); sw[j] = sin(T(2) * w * wts[j] * h); }
    T cw2 = cos(T(2) * w * h), sw2 = sin(T(2) * w * h);
    int cls = SIRIUS_RAY_MAX_STEPS;
    T rd = T(0), rprev = event_radius(s, q);
    uint32_t n = 0;
    while (n < (uint32_t)tp.max_steps) {
        T q0[4], p0[4];
        for (int i = 0; i < 4; ++i) { q0[i] = q[i]; p0[i] = p[i]; }
        const int subs = tp.sym_order >= 4 ? 3 : 1;
        for (int j = 0; j < subs; ++j) {
            T hj = subs == 3 ? wts[j] * h : h;
            T cj = subs == 3 ? cw[j] : cw2, sj = subs == 3 ? sw[j] : sw2;
            tao_a<T, FAM>(s, T(0.5) * hj, q, p, x, y);
            tao_b<T, FAM>(s, T(0.5) * hj, q, p, x, y);
            tao_c<T>(cj, sj, q, p, x, y);
            tao_b<T, FAM>(s, T(0.5) * hj, q, p, x, y);
            tao_a<T, FAM>(s, T(0.5) * hj, q, p, x, y);
        }
        n++;
        R.lam += h;
        bool bad = false;
        for (int i = 0; i < 4; ++i) bad |= !sfinite(q[i]) || !sfinite(p[i]);
        if (bad) { cls = SIRIUS_RAY_FAILED; break; }
        T r = event_radius(s, q);
        if (FAM == 0 && tp.emission == SIRIUS_EMISSION_THIN_DISK && q0[3] * q[3] < T(0)) {
            // the crossing is located on the cubic Hermite interpolant of the whole state over the
            // step (H's derivatives at both ends) and the ray ends there: a linear interpolant
            // would add an O(h^2) null-constraint error of its own at the hit
            Der<T> D0, D1;
            eval<T, FAM>(s, q0, p0, D0);
            eval<T, FAM>(s, q, p, D1);
            T th = locate_event(s, EV_Z, T(0), q0, D0.k, q, D1.k, h);
            T xc[4], pc[4];
            hermite(q0, D0.k, q, D1.k, h, th, xc);
            T rr = xc[1] * xc[1] + xc[2] * xc[2] - s.a2;
            T rdisk = sqrt(rr > T(0) ? rr : T(0));
            if (rdisk >= T(c.r_in) && rdisk <= T(c.r_out)) {
                const T f0[4] = {T(0), D0.f[0], D0.f[1], D0.f[2]}, f1[4] = {T(0), D1.f[0], D1.f[1], D1.f[2]};
                hermite(p0, f0, p, f1, h, th, pc);
                for (int i = 0; i < 4; ++i) { q[i] = xc[i]; p[i] = pc[i]; }
                R.lam += (th - T(1)) * h;
                rd = rdisk;
                cls = SIRIUS_RAY_DISK;
                break;
            }
        }
        if (FAM == 0 && r < T(c.r_cap)) { cls = SIRIUS_RAY_CAPTURED; break; }
        if (r > T(tp.r_escape) && r > rprev) { cls = SIRIUS_RAY_ESCAPED; break; }
        rprev = r;
    }
    for (int i = 0; i < 4; ++i) { R.x[i] = q[i]; R.p[i] = p[i]; }
    for (int j = 0; j < 8; ++j) R.lo[j] = T(0);  // the symplectic state is not compensated
    eval<T, FAM>(s, R.x, R.p, R.d);
    R.attempts = n;
    R.cls = cls;
    R.rdisk = rd;
    ray_store<T, FAM, SIRIUS_EMISSION_NONE>(R, o, tp);
}

}  // namespace sirius

// ------------------------------------------------------------------------------------------
// Host-callable launchers (C++ linkage, used by sirius_hip_api.cpp)
// ------------------------------------------------------------------------------------------
namespace sirius {

// Kernel variant for (precision, family, emission): emission only applies to Kerr-Schild.
#define SIRIUS_VARIANTS_M(MACRO, T, M)                                                            \
    switch (family) {                                                                             \
        case 1: MACRO(T, 1, SIRIUS_EMISSION_NONE, M); break;                                      \
        case 2: MACRO(T, 2, SIRIUS_EMISSION_NONE, M); break;                                      \
        default:                                                                                  \
            if (emission == SIRIUS_EMISSION_THIN_DISK) MACRO(T, 0, SIRIUS_EMISSION_THIN_DISK, M);    \
            else if (emission == SIRIUS_EMISSION_VOLUMETRIC) MACRO(T, 0, SIRIUS_EMISSION_VOLUMETRIC, M); \
            else MACRO(T, 0, SIRIUS_EMISSION_NONE, M);                                            \
            break;                                                                                \
    }
#define SIRIUS_VARIANTS(MACRO, T)                      \
    if (multi) { SIRIUS_VARIANTS_M(MACRO, T, true) }   \
    else { SIRIUS_VARIANTS_M(MACRO, T, false) }

int persistent_blocks_per_cu(bool fp64, int family, int emission, bool multi, int block) {
    int n = 0;
    hipError_t e = hipSuccess;
#define SIRIUS_OCC(T, F, E, M) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent<T, F, E, M>, block, 0)
#define SIRIUS_OCC32(T, F, E, M) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent32<F, E, M>, block, 0)
    if (fp64) { SIRIUS_VARIANTS(SIRIUS_OCC, double) } else { SIRIUS_VARIANTS(SIRIUS_OCC32, float) }
#undef SIRIUS_OCC32
#undef SIRIUS_OCC
    if (e != hipSuccess || n <= 0) n = 1;
    return n;
}

hipError_t launch_finish(int family, int emission, const CamConst* cams, const TraceParams& tp, const DevOut& o,
                         hipStream_t st) {
    dim3 g((unsigned)((o.n_rays + 255) / 256)), b(256);
    switch (family) {
        case 1: hipLaunchKernelGGL((k_finish<1, SIRIUS_EMISSION_NONE>), g, b, 0, st, cams, tp, o); break;
        case 2: hipLaunchKernelGGL((k_finish<2, SIRIUS_EMISSION_NONE>), g, b, 0, st, cams, tp, o); break;
        default:
            if (emission == SIRIUS_EMISSION_THIN_DISK)
                hipLaunchKernelGGL((k_finish<0, SIRIUS_EMISSION_THIN_DISK>), g, b, 0, st, cams, tp, o);
            else if (emission == SIRIUS_EMISSION_VOLUMETRIC)
                hipLaunchKernelGGL((k_finish<0, SIRIUS_EMISSION_VOLUMETRIC>), g, b, 0, st, cams, tp, o);
            else
                hipLaunchKernelGGL((k_finish<0, SIRIUS_EMISSION_NONE>), g, b, 0, st, cams, tp, o);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t* p, int n, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_trace_persistent(bool fp64, int family, int emission, bool multi, const CamConst* cams,
                                   const TraceParams& tp,
                                   const DevOut& o, const int32_t* list, const uint32_t* list_count, int64_t n,
                                   uint32_t* head, uint32_t* esc_count, int32_t* esc_list,
                                   unsigned long long* attempts_out, int grid, int block, int refill_min,
                                   uint32_t flag_or, hipStream_t st) {
    dim3 g(grid), b(block);
#define SIRIUS_PERS(T, F, E, M)                                                                               \
    hipLaunchKernelGGL((k_trace_persistent<T, F, E, M>), g, b, 0, st, cams, tp, o, list, list_count, n, head,    \
                       esc_count, esc_list, attempts_out, refill_min, flag_or)
#define SIRIUS_PERS32(T, F, E, M)                                                                             \
    hipLaunchKernelGGL((k_trace_persistent32<F, E, M>), g, b, 0, st, cams, tp, o, list, list_count, n, head,     \
                       esc_count, esc_list, attempts_out, refill_min, flag_or)
    if (fp64) { SIRIUS_VARIANTS(SIRIUS_PERS, double) } else { SIRIUS_VARIANTS(SIRIUS_PERS32, float) }
#undef SIRIUS_PERS32
#undef SIRIUS_PERS
    return hipGetLastError();
}

hipError_t launch_symplectic(bool fp64, int family, const CamConst* cams, const TraceParams& tp, const DevOut& o,
                             int block, hipStream_t st) {
    dim3 g((unsigned)((tp.n_rays + block - 1) / block)), b(block);
    if (fp64) {
        switch (family) {
            case 1: hipLaunchKernelGGL((k_symplectic_grid<double, 1>), g, b, 0, st, cams, tp, o); break;
            case 2: hipLaunchKernelGGL((k_symplectic_grid<double, 2>), g, b, 0, st, cams, tp, o); break;
            default: hipLaunchKernelGGL((k_symplectic_grid<double, 0>), g, b, 0, st, cams, tp, o); break;
        }
    } else {
        switch (family) {
            case 1: hipLaunchKernelGGL((k_symplectic_grid<float, 1>), g, b, 0, st, cams, tp, o); break;
            case 2: hipLaunchKernelGGL((k_symplectic_grid<float, 2>), g, b, 0, st, cams, tp, o); break;
            default: hipLaunchKernelGGL((k_symplectic_grid<float, 0>), g, b, 0, st, cams, tp, o); break;
        }
    }
    return hipGetLastError();
}

hipError_t launch_precision_select(const DevOut& o, int64_t n, double fraction, uint64_t seed, int32_t* list,
                                   uint32_t* count, hipStream_t st) {
    const uint64_t thr = fraction >= 1.0 ? ~0ull : (uint64_t)(fraction * 18446744073709551615.0);
    const int64_t per_block = 4 * 64 * (int64_t)PREC_SELECT_SPAN;  // four waves of 256 threads
    hipLaunchKernelGGL(k_precision_select, dim3((unsigned)((n + per_block - 1) / per_block)), dim3(256), 0, st, o, n,
                       thr, seed, list, count);
    return hipGetLastError();
}
hipError_t launch_precision_compare(const CamConst* cams, const TraceParams& tp, const DevOut& oa, const DevOut& o,
                                    const int32_t* list, const uint32_t* count, int64_t cap, unsigned int* acc,
                                    unsigned int* hist, hipStream_t st) {
    const int64_t blocks = std::min<int64_t>((cap + 255) / 256, PREC_COMPARE_BLOCKS);
    hipLaunchKernelGGL(k_precision_compare, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, st, cams, tp,
                       oa, o, list, count, cap, acc, hist);
    return hipGetLastError();
}

hipError_t launch_sur_audit(const CamConst* cams, const TraceParams& tp, const DevOut& oa, const DevOut& o,
                            const int32_t* list, const uint32_t* count, int64_t cap, unsigned int* acc,
                            hipStream_t st) {
    hipLaunchKernelGGL(k_sur_audit, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, st, cams, tp, oa, o, list,
                       count, cap, acc);
    return hipGetLastError();
}

}  // namespace sirius
