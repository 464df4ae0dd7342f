// trace_kernels.hip — geodesic tracing kernels for gfx950 (MI355X).
//
// Role replaced: the OptiX raygen mega-kernel's integrate/terminate/shade loop
// (src/Sirius.Render/Acceleration/OptiX/RDOP002A.cu:5132-6106), with the physics contract of
// the CPU tracer (src/Sirius.Render/Integration/GTRC001A.cpp:139-592) instead of the
// OptiX kernel's approximations. Three launch shapes:
//   k_trace_grid        one lane per ray, run to completion (plain FP64 path, escalations)
//   k_trace_persistent  a device-filling grid of waves that pull rays from an atomic work
//                       counter and refill idle lanes as rays terminate (FP32 bulk path);
//                       rays whose periapsis falls inside the photon region are pushed to an
//                       escalation queue and re-traced by k_trace_queue in FP64
//   k_symplectic_grid   fixed-step Tao/Yoshida integrator
#include <hip/hip_runtime.h>

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

// Null drift |g_{mu nu} k^mu k^nu| of the final state, evaluated in FP64.
template <int FAM>
SIRIUS_HD inline double null_drift64(const CamConst& c, const double x[4], const double p[4]) {
    Spacetime<double> s = make_spacetime<double>(c);
    Der<double> d;
    eval<double, FAM>(s, x, p, d);
    return fabs(d.pk);
}

// FP64 re-imposition of the null constraint on an FP32 state, keeping E = -p_t and the
// direction of the spatial momentum: p_i -> sigma p_i with H(sigma) = 0.
template <int FAM>
SIRIUS_HD inline void renorm_energy64(const Spacetime<double>& s, const float xf[4], float pf[4]) {
    double x[4] = {xf[0], xf[1], xf[2], xf[3]};
    double q[4] = {0.0, pf[1], pf[2], pf[3]};
    double e[4] = {pf[0], 0.0, 0.0, 0.0};
    Der<double> dq, de;
    eval<double, FAM>(s, x, q, dq);
    eval<double, FAM>(s, x, e, de);
    double A = q[1] * dq.k[1] + q[2] * dq.k[2] + q[3] * dq.k[3];
    double B = 2.0 * e[0] * dq.k[0];
    double C = e[0] * de.k[0];
    double disc = B * B - 4.0 * A * C;
    if (A <= 0.0 || disc < 0.0) return;
    double sq = sqrt(disc);
    double s1 = (-B + sq) / (2.0 * A), s2 = (-B - sq) / (2.0 * A);
    double sg = fabs(s1 - 1.0) < fabs(s2 - 1.0) ? s1 : s2;
    if (!(sg > 0.5 && sg < 2.0)) return;
    pf[1] = (float)(q[1] * sg);
    pf[2] = (float)(q[2] * sg);
    pf[3] = (float)(q[3] * sg);
}

// ------------------------------------------------------------------------------------------
// Per-lane ray state and one step attempt.
// ------------------------------------------------------------------------------------------
template <class T, int FAM>
struct Ray {
    T x[4], p[4];
    Der<T> d;
    T h, lam;
    T tau;          // volumetric optical depth
    T vr_prev;      // radial velocity sign tracking (FP32 escalation)
    int64_t id;
    int cam;
    int px, py;
    uint32_t attempts, rejected, landings, flags;
};

template <class T, int FAM>
SIRIUS_HD inline void ray_start(Ray<T, FAM>& R, const CamConst* cams, int cam, int64_t id, const TraceParams& tp,
                                 const DevOut& o) {
    const CamConst& c = cams[cam];
    int64_t pix = id - c.offset;
    R.id = id;
    R.cam = cam;
    R.px = (int)(pix % c.width);
    R.py = (int)(pix / c.width);
    Spacetime<T> s = make_spacetime<T>(c);
    init_state<T>(c, s, R.px, R.py, R.x, R.p);
    eval<T, FAM>(s, R.x, R.p, R.d);
    R.h = T(tp.h_init);
    R.lam = T(0);
    R.tau = T(0);
    R.vr_prev = T(-1);
    R.attempts = R.rejected = R.landings = R.flags = 0;
    if (tp.emission == SIRIUS_EMISSION_VOLUMETRIC)
        for (int k = 0; k < tp.n_bins; ++k) o.intensity[(int64_t)k * o.n_rays + id] = 0.f;
}

template <class T, int FAM>
SIRIUS_HD inline void ray_finish(Ray<T, FAM>& R, const CamConst* cams, const TraceParams& tp, const DevOut& o,
                                  int cls, T disk_r) {
    const CamConst& c = cams[R.cam];
    int64_t id = R.id;
    double xd[4], pd[4];
    for (int i = 0; i < 4; ++i) { xd[i] = (double)R.x[i]; pd[i] = (double)R.p[i]; }
    float th = 0.f, ph = 0.f, g = 0.f;
    if (cls == SIRIUS_RAY_ESCAPED) {
        double dir[3];
        if (FAM == 1) {
            // wormhole: map the orbital-plane end state back to the exit side's sky
            T ns[3];
            camera_dir<T>(c, R.px, R.py, ns);
            double nt = sqrt((double)ns[1] * ns[1] + (double)ns[2] * ns[2]);
            double e2[3];
            for (int i = 0; i < 3; ++i)
                e2[i] = nt > 1e-300 ? ((double)ns[1] * c.et[i] + (double)ns[2] * c.ep[i]) / nt : c.et[i];
            double l = xd[1], psi = xd[2], r = sqrt(l * l + c.b0 * c.b0);
            double drdl = (l / r) * (double)R.d.k[1], rdpsi = pd[2] / r;
            double cp = cos(psi), sp = sin(psi);
            for (int i = 0; i < 3; ++i) {
                double e = cp * c.er[i] + sp * e2[i], t = -sp * c.er[i] + cp * e2[i];
                dir[i] = drdl * e + rdpsi * t;
            }
            if (l < 0) R.flags |= SIRIUS_FLAG_FAR_SIDE;
        } else {
            dir[0] = (double)R.d.k[1];
            dir[1] = (double)R.d.k[2];
            dir[2] = (double)R.d.k[3];
        }
        double n = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        th = (float)acos(fmin(1.0, fmax(-1.0, dir[2] / n)));
        ph = (float)atan2(dir[1], dir[0]);
    }
    if (cls == SIRIUS_RAY_DISK && tp.emission == SIRIUS_EMISSION_THIN_DISK) {
        double gf;
        disk_emission(c, tp, xd, pd, (double)disk_r, gf, o.intensity + id, (int)o.n_rays);
        g = (float)gf;
    } else if (tp.emission == SIRIUS_EMISSION_VOLUMETRIC) {
        g = (float)R.tau;
    } else {
        for (int k = 0; k < tp.n_bins; ++k) o.intensity[(int64_t)k * o.n_rays + id] = 0.f;
    }
    o.cls[id] = (uint8_t)cls;
    o.flags[id] = (uint8_t)R.flags;
    o.theta[id] = th;
    o.phi[id] = ph;
    o.lam[id] = (float)R.lam;
    o.disk_r[id] = (float)disk_r;
    o.redshift[id] = g;
    o.null_drift[id] = (float)null_drift64<FAM>(c, xd, pd);
    o.steps[id] = R.attempts;
    o.rejected[id] = R.rejected;
    o.landings[id] = R.landings;
    if (o.state)
        for (int i = 0; i < 4; ++i) { o.state[8 * id + i] = xd[i]; o.state[8 * id + 4 + i] = pd[i]; }
}

// Volumetric transfer over an accepted step (3-node Gauss-Legendre on the Hermite interpolant
// of x, linear p). Emitter-frame j_nu = rho S(x_e), alpha = kappa rho, invariant form.
template <class T, int FAM>
SIRIUS_HD inline void vol_accumulate(Ray<T, FAM>& R, const CamConst& c, const Spacetime<T>& s, const TraceParams& tp,
                                      const DevOut& o, const T x1[4], const T p1[4], const Der<T>& d1) {
    const T gn[3] = {T(0.1127016653792583), T(0.5), T(0.8872983346207417)};
    const T gw[3] = {T(5.0 / 18.0), T(8.0 / 18.0), T(5.0 / 18.0)};
    Spacetime<double> sd = make_spacetime<double>(c);
    for (int q = 0; q < 3; ++q) {
        T xq[4];
        hermite(R.x, R.d.k, x1, d1.k, R.h, gn[q], xq);
        T r = kerr_r(s, xq[1], xq[2], xq[3]);
        if (r < T(tp.vol_r_min) || r > T(tp.vol_r_max)) continue;
        T sig = T(tp.vol_h_over_r) * r;
        T rho = T(tp.vol_j0) / (r * r) * exp(-xq[3] * xq[3] / (T(2) * sig * sig));
        // emitter velocity and frequencies in FP64
        double xd[4] = {(double)xq[0], (double)xq[1], (double)xq[2], (double)xq[3]};
        double pd[4];
        for (int i = 0; i < 4; ++i) pd[i] = (double)R.p[i] + (double)gn[q] * ((double)p1[i] - (double)R.p[i]);
        double rc = sqrt(fmax(xd[1] * xd[1] + xd[2] * xd[2] - c.a * c.a, 1e-12));
        double Om = kepler_omega(c, rc);
        double ut = emitter_ut(c, sd, xd, Om);
        if (ut < 0) continue;
        double nue = -ut * (pd[0] + Om * (xd[1] * pd[2] - xd[2] * pd[1]));
        double nuo = -pd[0] * c.obs_norm;
        if (nue <= 0 || nuo <= 0) continue;
        double gf = nuo / nue;
        double dl = (double)gw[q] * (double)R.h;
        double dtau = tp.vol_kappa * (double)rho * nue * dl;
        double w = (double)rho * (nuo * nuo * nuo / (nue * nue)) * dl * exp(-((double)R.tau + 0.5 * dtau));
        double tv = pow(6.0 / (double)r, 0.75);
        double lr = log(tp.x_max / tp.x_min);
        for (int k = 0; k < tp.n_bins; ++k) {
            double xk = tp.x_min * exp(lr * (k + 0.5) / tp.n_bins);
            double xe = xk / gf;
            o.intensity[(int64_t)k * o.n_rays + R.id] += (float)(w * xe * xe * xe / expm1(xe / tv));
        }
        R.tau += T(dtau);
    }
}

template <class T>
SIRIUS_HD __forceinline__ T vol_step_cap(const Spacetime<T>& s, const TraceParams& tp, const T x[4]) {
    T r = kerr_r(s, x[1], x[2], x[3]);
    if (r < T(0.5 * tp.vol_r_min) || r > T(1.5 * tp.vol_r_max)) return T(1e30);
    T sig = T(tp.vol_h_over_r) * r;
    T cap = T(0.5) * (fabs(x[3]) - T(3) * sig);
    return cap > T(0.2) * sig ? cap : T(0.2) * sig;
}

// Land exactly on an event inside the step: Hermite bisection, one DP45 step of th*h, and
// for the disk plane one Newton correction of th. Returns th; xl/pl/dl hold the landed state.
template <class T, int FAM>
SIRIUS_HD inline T land(Ray<T, FAM>& R, const Spacetime<T>& s, int kind, T target, const T x1[4], const Der<T>& d1,
                         T xl[4], T pl[4], Der<T>& dl) {
    T th = locate_event(s, kind, target, R.x, R.d.k, x1, d1.k, R.h);
    T err[4];
    dp45_step<T, FAM>(s, R.x, R.p, R.d, th * R.h, xl, pl, dl, err, false);
    R.landings++;
    if (kind == EV_Z && fabs(dl.k[3]) > T(1e-30)) {
        T th2 = th - xl[3] / (dl.k[3] * R.h);
        if (th2 > T(0) && th2 < T(1.5)) {
            th = th2;
            dp45_step<T, FAM>(s, R.x, R.p, R.d, th * R.h, xl, pl, dl, err, false);
            R.landings++;
        }
    }
    return th;
}

// One DP45 attempt with error control, events and acceptance. Returns ST_RUN, ST_DONE (ray
// finished and written) or ST_ESCALATE (FP32 lane hands the ray to the FP64 queue).
template <class T, int FAM>
SIRIUS_HD inline int ray_attempt(Ray<T, FAM>& R, const CamConst* cams, const Spacetime<T>& s, const TraceParams& tp,
                                  const DevOut& o) {
    const bool fp32 = sizeof(T) == 4;
    const bool vol = (FAM == 0) && tp.emission == SIRIUS_EMISSION_VOLUMETRIC;
    T hmax = T(tp.h_max);
    if (fp32 && tp.far_slope > 0) {
        T rr = event_radius(s, R.x) * T(tp.far_slope);
        hmax = rr > hmax ? rr : hmax;
    }
    if (vol) {
        T cap = vol_step_cap(s, tp, R.x);
        if (R.h > cap) R.h = cap;
    }
    T x1[4], p1[4], err[4];
    Der<T> d1;
    dp45_step<T, FAM>(s, R.x, R.p, R.d, R.h, x1, p1, d1, err, true);
    R.attempts++;
    T e = T(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T q = err[i] / (T(tp.atol) + T(tp.rtol) * fabs(x1[i]));
        e += q * q;
    }
    e = sqrt(e * T(0.25));
    if (!(e <= T(1))) {
        if (!sfinite(e)) { ray_finish(R, cams, tp, o, SIRIUS_RAY_FAILED, T(0)); return ST_DONE; }
        R.h = next_step(R.h, e, T(tp.h_min), hmax);
        R.rejected++;
        if (R.h <= T(tp.h_min)) { ray_finish(R, cams, tp, o, SIRIUS_RAY_FAILED, T(0)); return ST_DONE; }
        if (R.attempts >= (uint32_t)tp.max_steps) { ray_finish(R, cams, tp, o, SIRIUS_RAY_MAX_STEPS, T(0)); return ST_DONE; }
        return ST_RUN;
    }
    bool bad = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) bad |= !sfinite(x1[i]) || !sfinite(p1[i]);
    if (bad) { ray_finish(R, cams, tp, o, SIRIUS_RAY_FAILED, T(0)); return ST_DONE; }
    const CamConst& c = cams[R.cam];
    // ---- volumetric transfer ----
    if (vol) {
        vol_accumulate<T, FAM>(R, c, s, tp, o, x1, p1, d1);
        if (R.tau > T(tp.vol_tau_max)) {
            R.lam += R.h;
            for (int i = 0; i < 4; ++i) { R.x[i] = x1[i]; R.p[i] = p1[i]; }
            R.d = d1;
            ray_finish(R, cams, tp, o, SIRIUS_RAY_DISK, T(0));
            return ST_DONE;
        }
    }
    // ---- thin disk: first crossing of the equatorial plane inside [r_in, r_out] ----
    if (FAM == 0 && tp.emission == SIRIUS_EMISSION_THIN_DISK && R.x[3] * x1[3] < T(0)) {
        T xl[4], pl[4];
        Der<T> dl;
        T th = land<T, FAM>(R, s, EV_Z, T(0), x1, d1, xl, pl, dl);
        T rd = kerr_r(s, xl[1], xl[2], xl[3]);
        if (rd >= T(c.r_in) && rd <= T(c.r_out)) {
            R.lam += th * R.h;
            for (int i = 0; i < 4; ++i) { R.x[i] = xl[i]; R.p[i] = pl[i]; }
            R.d = dl;
            ray_finish(R, cams, tp, o, SIRIUS_RAY_DISK, rd);
            return ST_DONE;
        }
    }
    T r1 = event_radius(s, x1);
    // ---- horizon ----
    if (FAM == 0 && r1 < T(c.r_cap)) {
        T th = T(1);
        if (event_radius(s, R.x) > T(c.r_cap)) {
            T xl[4], pl[4];
            Der<T> dl;
            th = land<T, FAM>(R, s, EV_R, T(c.r_cap), x1, d1, xl, pl, dl);
            for (int i = 0; i < 4; ++i) { x1[i] = xl[i]; p1[i] = pl[i]; }
            d1 = dl;
        }
        R.lam += th * R.h;
        for (int i = 0; i < 4; ++i) { R.x[i] = x1[i]; R.p[i] = p1[i]; }
        R.d = d1;
        ray_finish(R, cams, tp, o, SIRIUS_RAY_CAPTURED, T(0));
        return ST_DONE;
    }
    // ---- escape sphere ----
    if (r1 > T(tp.r_escape)) {
        T radial = (FAM == 1) ? x1[1] * d1.k[1] : x1[1] * d1.k[1] + x1[2] * d1.k[2] + x1[3] * d1.k[3];
        if (radial > T(0)) {
            T th = T(1);
            if (event_radius(s, R.x) < T(tp.r_escape)) {
                T xl[4], pl[4];
                Der<T> dl;
                th = land<T, FAM>(R, s, EV_R, T(tp.r_escape), x1, d1, xl, pl, dl);
                for (int i = 0; i < 4; ++i) { x1[i] = xl[i]; p1[i] = pl[i]; }
                d1 = dl;
            }
            R.lam += th * R.h;
            for (int i = 0; i < 4; ++i) { R.x[i] = x1[i]; R.p[i] = p1[i]; }
            R.d = d1;
            ray_finish(R, cams, tp, o, SIRIUS_RAY_ESCAPED, T(0));
            return ST_DONE;
        }
    }
    // ---- accept ----
    R.lam += R.h;
    T hmax2 = T(tp.h_max);
    if (fp32 && tp.far_slope > 0) {
        T rr = r1 * T(tp.far_slope);
        hmax2 = rr > hmax2 ? rr : hmax2;
    }
    R.h = next_step(R.h, e, T(tp.h_min), hmax2);
    for (int i = 0; i < 4; ++i) { R.x[i] = x1[i]; R.p[i] = p1[i]; }
    R.d = d1;
    if constexpr (sizeof(T) == 4) {
        // escalation: a radial turning point inside the photon region
        if (FAM == 0 && tp.escalate) {
            T vr = x1[1] * d1.k[1] + x1[2] * d1.k[2] + x1[3] * d1.k[3];
            if (R.vr_prev < T(0) && vr >= T(0) && r1 < T(tp.escalate_margin * c.r_ph_retro)) return ST_ESCALATE;
            R.vr_prev = vr;
        }
        // FP64 re-imposition of the null constraint, energy preserving
        if (fabs(R.d.pk) > T(1e-6) * R.p[0] * R.p[0]) {
            Spacetime<double> sd = make_spacetime<double>(c);
            renorm_energy64<FAM>(sd, R.x, R.p);
            eval<T, FAM>(s, R.x, R.p, R.d);
            R.flags |= SIRIUS_FLAG_RENORM;
        }
    } else if (fabs(R.d.pk) > T(1e-6)) {
        // FP64 path: reference re-normalisation: re-solve k^0 with k^i fixed (PHGD001A.cpp:609-638)
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
    if (R.attempts >= (uint32_t)tp.max_steps) { ray_finish(R, cams, tp, o, SIRIUS_RAY_MAX_STEPS, T(0)); return ST_DONE; }
    return ST_RUN;
}

// ------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------

// One lane per ray, run to completion. ray_list (optional) maps lane -> ray; count_dev
// (optional) holds the list length on the device.
template <class T, int FAM>
__global__ __launch_bounds__(256) void k_trace_grid(const CamConst* __restrict__ cams, TraceParams tp, DevOut o,
                                                    const int32_t* __restrict__ ray_list, int64_t n,
                                                    const uint32_t* __restrict__ count_dev, uint32_t flag_or) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t lim = count_dev ? (int64_t)*count_dev : n;
    if (lim > n) lim = n;
    if (i >= lim) return;
    int64_t id = ray_list ? (int64_t)ray_list[i] : i;
    int cam = tp.n_cam == 1 ? 0 : find_camera(cams, tp.n_cam, id);
    Ray<T, FAM> R;
    ray_start<T, FAM>(R, cams, cam, id, tp, o);
    R.flags |= flag_or;
    Spacetime<T> s = make_spacetime<T>(cams[cam]);
    while (true) {
        int st = ray_attempt<T, FAM>(R, cams, s, tp, o);
        if (st != ST_RUN) break;
    }
}

// Persistent tracer. Each wave pulls batches of work-list slots from *head (one atomic per
// refill) and refills its idle lanes when at least refill_min of them are idle; slot -> ray
// through ray_list when given. FP32 lanes that must escalate append their ray to esc_list
// (count in *esc_count) and add nothing else; each lane adds the attempts of every ray it
// traced to *attempts_out when the wave retires.
template <class T, int FAM>
__global__ __launch_bounds__(256) void k_trace_persistent(const CamConst* __restrict__ cams, TraceParams tp, DevOut o,
                                                          const int32_t* __restrict__ ray_list,
                                                          const uint32_t* __restrict__ list_count, int64_t n,
                                                          uint32_t* __restrict__ head, uint32_t* __restrict__ esc_count,
                                                          int32_t* __restrict__ esc_list,
                                                          unsigned long long* __restrict__ attempts_out,
                                                          int refill_min, uint32_t flag_or) {
    const int lane = __lane_id();
    int64_t total = list_count ? (int64_t)*list_count : n;
    if (total > n) total = n;
    Ray<T, FAM> R;
    Spacetime<T> s;
    bool active = false;
    bool exhausted = false;
    unsigned long long my_attempts = 0;  // attempts of every ray this lane traced
    while (true) {
        unsigned long long idle = __ballot(!active);
        int n_idle = __popcll(idle);
        if (n_idle == 64 && exhausted) break;
        if (!exhausted && (n_idle >= refill_min || n_idle == 64)) {
            int leader = __ffsll((long long)idle) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(head, (uint32_t)n_idle);
            base = __shfl(base, leader);
            if ((int64_t)base + n_idle >= total) exhausted = true;
            if (!active) {
                unsigned long long below = lane ? (idle & ((1ull << lane) - 1ull)) : 0ull;
                int64_t slot = (int64_t)base + __popcll(below);
                if (slot < total) {
                    int64_t id = ray_list ? (int64_t)ray_list[slot] : slot;
                    int cam = tp.n_cam == 1 ? 0 : find_camera(cams, tp.n_cam, id);
                    ray_start<T, FAM>(R, cams, cam, id, tp, o);
                    R.flags |= flag_or;
                    s = make_spacetime<T>(cams[cam]);
                    active = true;
                }
            }
        }
        if (active) {
            int st = ray_attempt<T, FAM>(R, cams, s, tp, o);
            if (st == ST_ESCALATE) {
                uint32_t k = atomicAdd(esc_count, 1u);
                esc_list[k] = (int32_t)R.id;
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
SIRIUS_HD inline void tao_a(const Spacetime<T>& s, T d, T q[4], T p[4], T x[4], const T y[4]) {
    Der<T> D;
    eval<T, FAM>(s, q, y, D);
    for (int i = 0; i < 3; ++i) p[i + 1] += d * D.f[i];
    for (int i = 0; i < 4; ++i) x[i] += d * D.k[i];
}
template <class T, int FAM>
SIRIUS_HD inline void tao_b(const Spacetime<T>& s, T d, T q[4], const T p[4], const T x[4], T y[4]) {
    Der<T> D;
    eval<T, FAM>(s, x, p, D);
    for (int i = 0; i < 4; ++i) q[i] += d * D.k[i];
    for (int i = 0; i < 3; ++i) y[i + 1] += d * D.f[i];
}
template <class T>
SIRIUS_HD inline void tao_c(T cw, T sw, T q[4], T p[4], T x[4], T y[4]) {
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
    ray_start<T, FAM>(R, cams, cam, id, tp, o);
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
            T al = -q0[3] / (q[3] - q0[3]);
            T xc = q0[1] + al * (q[1] - q0[1]), yc = q0[2] + al * (q[2] - q0[2]);
            T rr = xc * xc + yc * yc - s.a2;
            T rdisk = sqrt(rr > T(0) ? rr : T(0));
            if (rdisk >= T(c.r_in) && rdisk <= T(c.r_out)) {
                for (int i = 0; i < 4; ++i) { q[i] = q0[i] + al * (q[i] - q0[i]); p[i] = p0[i] + al * (p[i] - p0[i]); }
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
    eval<T, FAM>(s, R.x, R.p, R.d);
    R.attempts = n;
    ray_finish(R, cams, tp, o, cls, rd);
}

}  // namespace sirius

// ------------------------------------------------------------------------------------------
// Host-callable launchers (C++ linkage, used by sirius_hip_api.cpp)
// ------------------------------------------------------------------------------------------
namespace sirius {

int persistent_blocks_per_cu(bool fp64, int family, int block) {
    int n = 0;
    hipError_t e;
    if (fp64) {
        switch (family) {
            case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent<double, 1>, block, 0); break;
            case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent<double, 2>, block, 0); break;
            default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent<double, 0>, block, 0); break;
        }
    } else {
        switch (family) {
            case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent<float, 1>, block, 0); break;
            case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent<float, 2>, block, 0); break;
            default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_persistent<float, 0>, block, 0); break;
        }
    }
    if (e != hipSuccess || n <= 0) n = 1;
    return n;
}

hipError_t launch_fill(uint32_t* p, int n, uint32_t v, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

template <class T>
static hipError_t grid_dispatch(int family, const CamConst* cams, const TraceParams& tp, const DevOut& o,
                                const int32_t* list, int64_t n, const uint32_t* count_dev, uint32_t flag_or, int block,
                                hipStream_t st) {
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + block - 1) / block)), b(block);
    switch (family) {
        case 1: hipLaunchKernelGGL((k_trace_grid<T, 1>), g, b, 0, st, cams, tp, o, list, n, count_dev, flag_or); break;
        case 2: hipLaunchKernelGGL((k_trace_grid<T, 2>), g, b, 0, st, cams, tp, o, list, n, count_dev, flag_or); break;
        default: hipLaunchKernelGGL((k_trace_grid<T, 0>), g, b, 0, st, cams, tp, o, list, n, count_dev, flag_or); break;
    }
    return hipGetLastError();
}

hipError_t launch_trace_grid(bool fp64, int family, const CamConst* cams, const TraceParams& tp, const DevOut& o,
                             const int32_t* list, int64_t n, const uint32_t* count_dev, uint32_t flag_or, int block,
                             hipStream_t st) {
    return fp64 ? grid_dispatch<double>(family, cams, tp, o, list, n, count_dev, flag_or, block, st)
                : grid_dispatch<float>(family, cams, tp, o, list, n, count_dev, flag_or, block, st);
}

hipError_t launch_trace_persistent(bool fp64, int family, const CamConst* cams, const TraceParams& tp,
                                   const DevOut& o, const int32_t* list, const uint32_t* list_count, int64_t n,
                                   uint32_t* head, uint32_t* esc_count, int32_t* esc_list,
                                   unsigned long long* attempts_out, int grid, int block, int refill_min,
                                   uint32_t flag_or, hipStream_t st) {
    dim3 g(grid), b(block);
#define SIRIUS_LAUNCH_P(T, F)                                                                                   \
    hipLaunchKernelGGL((k_trace_persistent<T, F>), g, b, 0, st, cams, tp, o, list, list_count, n, head, esc_count, \
                       esc_list, attempts_out, refill_min, flag_or)
    if (fp64) {
        switch (family) {
            case 1: SIRIUS_LAUNCH_P(double, 1); break;
            case 2: SIRIUS_LAUNCH_P(double, 2); break;
            default: SIRIUS_LAUNCH_P(double, 0); break;
        }
    } else {
        switch (family) {
            case 1: SIRIUS_LAUNCH_P(float, 1); break;
            case 2: SIRIUS_LAUNCH_P(float, 2); break;
            default: SIRIUS_LAUNCH_P(float, 0); break;
        }
    }
#undef SIRIUS_LAUNCH_P
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

}  // namespace sirius
