// geodesic.hpp — device-side physics of the MI355X geodesic backend, templated on the scalar
// type (double for the FP64 path, float for the FP32 bulk path).
//
// One lane owns one ray; every quantity below lives in VGPRs. Reference lines each part
// reproduces (paths relative to the reference's src/):
//   Kerr-Schild metric and gradient  Sirius.Core/Metric/PHMT100A.h:248-396
//   Ellis wormhole                   Sirius.Core/Metric/PHMT101A.h:285-345 (two-sided l form)
//   Alcubierre                       Sirius.Core/Metric/PHMT102A.h:485-578 (static bubble)
//   Hamiltonian RHS                  Sirius.Core/Geodesic/PHGD001A.cpp:159-202
//   DP45 tableau / step control      Sirius.Core/Geodesic/PHGD001A.cpp:16-79,471-583
//   null re-normalisation            Sirius.Core/Tensor/MTTN001A.cpp:171-237
//   camera                           Sirius.Core/Camera/CMBS001A.h:120-185
//   initial state                    Sirius.Render/Integration/GTRC001A.cpp:15-134
// The metric gradient comes from forward-mode dual numbers with three tangents (the role of
// Dual<T>, Sirius.Core/Autodiff/MTDL001A.h:19-368), and g^{mu nu} from the Kerr-Schild
// closed form eta - H l l (plus a Sherman-Morrison rank-1 term for Lambda), so no 4x4
// inverse is ever formed on the device.
#pragma once
#include <cmath>
#include <type_traits>

#include <hip/hip_runtime.h>
#include <stdint.h>

// Every device function is also a host function, so the identical per-ray code can be
// compiled for the CPU by diagnostic harnesses.
#define SIRIUS_HD __host__ __device__

namespace sirius {

// ------------------------------------------------------------------------------------------
// Per-camera constants, computed once on the host in FP64 (sirius_hip_api.cpp).
// ------------------------------------------------------------------------------------------
struct CamConst {
    // camera model
    double r_obs, theta_obs, phi_obs;
    double tan_half, aspect;
    double cyaw, syaw, cpitch, spitch, croll, sroll;
    double er[3], et[3], ep[3];      // spherical basis at the observer
    int32_t width, height, family, pad0;
    int64_t offset;                  // first global ray number of this camera
    // spacetime
    double M, a, Q2, Lambda;
    double b0;                       // wormhole throat
    double wv, wR, wsig, winv2tanh; // warp: v, R, sigma, 1/(2 tanh(sigma R))
    double wE, wN;                  // warp: e^{-4 sigma R}, (1 - e^{-4 sigma R}) / tanh(sigma R)
    // termination and emission
    double r_cap, r_in, r_out;
    double fmax;                     // max of f(r)/r over the disk (Page-Thorne)
    double obs_norm;                 // 1/sqrt(-g_tt) at the observer
    double r_ph_retro;               // retrograde circular photon orbit
    double pt_x0, pt_x[3], pt_c[3];  // Page-Thorne closed-form constants
    double l_obs;                    // wormhole: observer's proper-length coordinate
    double g_obs[16];                // g_{mu nu} at the observer (every ray's initial position)
};

struct TraceParams {
    double rtol, atol, h_init, h_min, h_max, far_slope;
    double r_escape;
    double x_min, x_max;
    double bin_x0, bin_q;  // spectral bin centres: the first, and the ratio of neighbours (host-computed)
    double escalate_margin;
    double sym_h;
    double vol_j0, vol_h_over_r, vol_kappa, vol_r_min, vol_r_max, vol_tau_max;
    double esc_edge;
    double step_rexp;  // inward step-proposal exponent (0: the reference's controller)
    int32_t max_steps, n_bins, emission, sym_order;
    int32_t n_cam, escalate, gate_flags;  // gate_flags: the gate wrote per-ray flags
    int32_t spectrum32;  // FP32 path: thin-disk Planck bins of FP32-traced rays in FP32 (g, T in FP64)
    int64_t n_rays;
    // persistent queue: rays listed here are taken before the kernel's own list (the rays
    // nearest the critical curve, the longest, start first); null for none
    const int32_t* first_list;
    const uint32_t* first_count;
};

// Null-drift histogram of the statistics pass: log10 bins of 1/16 decade over [1e-20, 1).
constexpr int SIRIUS_DRIFT_BINS = 320;

// Floats per device star record of a celestial sphere: unit direction (3), cosine cut,
// sigma^2 of the point-spread function, 10^(-0.4 m), colour (3), padding (3).
constexpr int SIRIUS_STAR_REC = 12;

// Finite test usable on host and device: x - x is NaN exactly when x is NaN or infinite.
template <class T>
SIRIUS_HD __forceinline__ bool sfinite(T x) {
    T y = x - x;
    return y == y;
}

// Framebuffer shading of launch (the public SiriusHipShading's fields the shading kernel reads).
struct ShadeParams {
    int32_t temperature_model;  // 0 Shakura-Sunyaev, 1 Novikov-Thorne
    double disk_temperature;    // K
    double disk_boost;
    double max_steps_level;
};

// Device result arrays (structure of arrays, one slot per ray; intensity is bin-major).
struct DevOut {
    uint8_t* cls;
    uint8_t* flags;
    float* theta;
    float* phi;
    float* lam;
    float* disk_r;
    float* redshift;
    float* null_drift;
    uint32_t* steps;
    uint32_t* rejected;
    uint32_t* landings;
    float* intensity;
    double* state;  // optional, 8 per ray
    double* vol_acc;  // volumetric emission: per-bin FP64 accumulators (bin-major), or null
    int64_t n_rays;
};

// Reciprocal square root and reciprocal: in FP32 device code the hardware instructions
// (about 1 ulp), elsewhere the correctly rounded forms.
template <class T> SIRIUS_HD __forceinline__ T rsqrt_t(T x) {
#ifdef __HIP_DEVICE_COMPILE__
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rsqf(x);
#endif
    return T(1) / sqrt(x);
}
template <class T> SIRIUS_HD __forceinline__ T rcp_t(T x) {
#ifdef __HIP_DEVICE_COMPILE__
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
#endif
    return T(1) / x;
}

// a * b + c with one rounding in the working precision (fmaf for FP32, also in host builds of
// this header, where a plain fma call on floats would resolve to the double overload).
template <class T> SIRIUS_HD __forceinline__ T tfma(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return fmaf(a, b, c);
    else return fma(a, b, c);
}

// ------------------------------------------------------------------------------------------
// Dual numbers (forward-mode derivatives, MTDL001A.h's Dual<T> with N tangents): a value and its
// derivatives along N seed directions, all held in registers. The geodesic code is written over
// a scalar type T; instantiated with Dual<double, N> it carries the derivatives of every quantity
// with respect to N parameters through the same arithmetic (the ray bundles differentiate the
// camera, the initial state and each DP45 step with respect to the image point this way).
// Comparisons and the clamps fmax / fmin act on values.
// ------------------------------------------------------------------------------------------
template <class T, int N>
struct Dual {
    T v;
    T d[N];
    Dual() = default;
    template <class S, std::enable_if_t<std::is_arithmetic_v<S>, int> = 0>
    SIRIUS_HD constexpr Dual(S x) : v(T(x)), d{} {}
    SIRIUS_HD static Dual seed(T x, int axis) {
        Dual r(x);
        r.d[axis] = T(1);
        return r;
    }
    SIRIUS_HD Dual& operator+=(const Dual& b) { *this = *this + b; return *this; }
    SIRIUS_HD Dual& operator-=(const Dual& b) { *this = *this - b; return *this; }
    SIRIUS_HD Dual& operator*=(const Dual& b) { *this = *this * b; return *this; }
    SIRIUS_HD Dual& operator/=(const Dual& b) { *this = *this / b; return *this; }
};
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> operator+(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> operator-(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> operator-(const Dual<T, N>& a) {
    Dual<T, N> r;
    r.v = -a.v;
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> operator*(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> operator/(const Dual<T, N>& a, const Dual<T, N>& b) {
    Dual<T, N> r;
    const T ib = T(1) / b.v;
    r.v = a.v * ib;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
template <class T, int N> SIRIUS_HD __forceinline__ bool operator<(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v < b.v; }
template <class T, int N> SIRIUS_HD __forceinline__ bool operator>(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v > b.v; }
template <class T, int N> SIRIUS_HD __forceinline__ bool operator<=(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v <= b.v; }
template <class T, int N> SIRIUS_HD __forceinline__ bool operator>=(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v >= b.v; }
template <class T, int N> SIRIUS_HD __forceinline__ bool operator==(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v == b.v; }
template <class T, int N> SIRIUS_HD __forceinline__ bool operator!=(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v != b.v; }
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> chain(const Dual<T, N>& a, T f, T df) {
    Dual<T, N> r;
    r.v = f;
    for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i];
    return r;
}
// the scalar overloads stay visible beside the Dual ones inside this namespace
using ::fabs;
using ::fma;
using ::fmax;
using ::fmin;
using ::exp;
using ::sqrt;
using ::tanh;
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> sqrt(const Dual<T, N>& a) {
    const T s = sqrt(a.v);
    return chain(a, s, T(0.5) / s);
}
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> fabs(const Dual<T, N>& a) { return a.v < T(0) ? -a : a; }
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> fmax(const Dual<T, N>& a, const Dual<T, N>& b) { return a.v < b.v ? b : a; }
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> fmin(const Dual<T, N>& a, const Dual<T, N>& b) { return b.v < a.v ? b : a; }
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> exp(const Dual<T, N>& a) {
    const T e = exp(a.v);
    return chain(a, e, e);
}
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> tanh(const Dual<T, N>& a) {
    const T t = tanh(a.v);
    return chain(a, t, T(1) - t * t);
}
template <class T, int N> SIRIUS_HD __forceinline__ Dual<T, N> fma(const Dual<T, N>& a, const Dual<T, N>& b, const Dual<T, N>& c) {
    Dual<T, N> r;
    r.v = fma(a.v, b.v, c.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i] + c.d[i];
    return r;
}

// Scalar parameters of a spacetime in the working precision.
template <class T>
struct Spacetime {
    int family;
    T M, a, a2, Q2, Lambda;
    T b0, b02;
    T wv, wR, wsig, winv2tanh;
    T wE, wN;
};

template <class T>
SIRIUS_HD __forceinline__ Spacetime<T> make_spacetime(const CamConst& c) {
    Spacetime<T> s;
    s.family = c.family;
    s.M = T(c.M); s.a = T(c.a); s.a2 = T(c.a * c.a); s.Q2 = T(c.Q2); s.Lambda = T(c.Lambda);
    s.b0 = T(c.b0); s.b02 = T(c.b0 * c.b0);
    s.wv = T(c.wv); s.wR = T(c.wR); s.wsig = T(c.wsig); s.winv2tanh = T(c.winv2tanh);
    s.wE = T(c.wE); s.wN = T(c.wN);
    return s;
}

// Alcubierre shape function f(r_s) = (tanh(sigma (r_s + R)) - tanh(sigma (r_s - R))) / (2 tanh(sigma R))
// and df/dr_s, by tanh A - tanh B = sinh(A - B) / (cosh A cosh B) with A - B = 2 sigma R: with
// v = e^{-2 sigma (r_s - R)}, E = e^{-4 sigma R} and N = (1 - E) / tanh(sigma R),
//   f = N v / D,  df/dr_s = -2 sigma f (1 - E v^2) / D,  D = 1 + v (1 + E + E v).
// One exponential instead of two hyperbolic tangents; its argument is small at the wall, where f
// varies, so it keeps the digits the tanh of sigma (r_s - R) keeps; E v = e^{-2 sigma (r_s + R)}
// and E v^2 = e^{-4 sigma r_s} stay below one, and v below e^{2 sigma R} (sigma R <= 40).
template <class T>
SIRIUS_HD __forceinline__ T warp_shape(const Spacetime<T>& s, T rs, T& dfdr) {
    const T v = exp(-T(2) * s.wsig * (rs - s.wR));
    const T ev = s.wE * v;
    const T D = tfma(v, T(1) + s.wE + ev, T(1));
    const T iD = T(1) / D;
    const T f = s.wN * v * iD;
    dfdr = -T(2) * s.wsig * f * (T(1) - ev * v) * iD;
    return f;
}

// Kerr radius (value only), PHMT100A.h:198-216.
template <class T>
SIRIUS_HD __forceinline__ T kerr_r(const Spacetime<T>& s, T x, T y, T z) {
    T R2 = x * x + y * y + z * z;
    if (s.family != 0) return sqrt(R2);
    T w = R2 - s.a2;
    T disc = w * w + T(4) * s.a2 * z * z;
    disc = disc < T(1e-20) ? T(1e-20) : disc;
    T r2 = T(0.5) * (w + sqrt(disc));
    r2 = r2 < T(1e-10) ? T(1e-10) : r2;
    return sqrt(r2);
}

// Event radius used for termination: Kerr r, Euclidean R (warp) or |l| (wormhole).
template <class T>
SIRIUS_HD __forceinline__ T event_radius(const Spacetime<T>& s, const T x[4]) {
    if (s.family == 1) return fabs(x[1]);
    return kerr_r(s, x[1], x[2], x[3]);
}

// Event radius and its rate of change along the velocity v (v[1..3]): for the Kerr r the
// implicit derivative of r^4 - (R^2 - a^2) r^2 - a^2 z^2 = 0, dr = (r (x dx + y dy) +
// (r^2 + a^2) z dz / r) / (2 r^2 - R^2 + a^2); R for the warp bubble, |l| for the wormhole.
template <class T>
SIRIUS_HD __forceinline__ T event_radius_rate(const Spacetime<T>& s, const T x[4], const T v[4], T& rate) {
    if (s.family == 1) {
        rate = x[1] < T(0) ? -v[1] : v[1];
        return fabs(x[1]);
    }
    const T R2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
    if (s.family != 0) {
        T R = sqrt(R2);
        rate = (x[1] * v[1] + x[2] * v[2] + x[3] * v[3]) / R;
        return R;
    }
    const T r = kerr_r(s, x[1], x[2], x[3]);
    const T r2 = r * r;
    const T den = T(2) * r2 - R2 + s.a2;
    rate = (r * (x[1] * v[1] + x[2] * v[2]) + (r2 + s.a2) * x[3] * v[3] / r) / den;
    return r;
}

// Covariant metric values (no gradient), for initial conditions and emitter velocities.
template <class T>
SIRIUS_HD __forceinline__ void metric_cov(const Spacetime<T>& s, const T x[4], T g[4][4]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g[i][j] = (i == j) ? (i == 0 ? T(-1) : T(1)) : T(0);
    if (s.family == 1) {
        g[2][2] = x[1] * x[1] + s.b02;
        return;
    }
    if (s.family == 2) {
        T rs = sqrt(x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
        T dfdr;
        T f = warp_shape(s, rs, dfdr);
        g[0][0] = -(T(1) - s.wv * s.wv * f * f);
        g[0][1] = g[1][0] = -s.wv * f;
        return;
    }
    T r = kerr_r(s, x[1], x[2], x[3]);
    T r2 = r * r, den = T(1) / (r2 + s.a2);
    T l[4] = {T(1), (r * x[1] + s.a * x[2]) * den, (r * x[2] - s.a * x[1]) * den, x[3] / r};
    T H = (T(2) * s.M * r - s.Q2) * r2 / (r2 * r2 + s.a2 * x[3] * x[3]);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g[i][j] += H * l[i] * l[j];
    if (s.Lambda != T(0)) g[0][0] -= s.Lambda * (x[1] * x[1] + x[2] * x[2] + x[3] * x[3]) / T(3);
}

// ------------------------------------------------------------------------------------------
// Hamiltonian right-hand side: k^mu = g^{mu nu} p_nu (written to dx), and
// dp_i = 1/2 d_i g_{rs} k^r k^s for i = 1..3 (dp_t = 0: all metrics are stationary).
// Returns p.k = g_{mu nu} k^mu k^nu (the null-constraint value).
// ------------------------------------------------------------------------------------------
template <class T>
SIRIUS_HD __forceinline__ T rhs_ks(const Spacetime<T>& s, const T x[4], const T p[4], T k[4], T dp[3]) {
    // Hamilton's equations for H = (eta^{mu nu} - H l^mu l^nu) p_mu p_nu / 2 with the position
    // gradient in closed form. r depends on (x, y, z) through r^4 - (R^2 - a^2) r^2 - a^2 z^2 = 0:
    //   dr/dx = x r / s, dr/dy = y r / s, dr/dz = z (r^2 + a^2) / (r s),
    //   s = sqrt((R^2 - a^2)^2 + 4 a^2 z^2) = 2 r^2 - R^2 + a^2;
    // then dH_i = H_r dr_i + H_z [i = z] and sum_j k_j d_i l_j = A_i + B dr_i.
    constexpr bool f32 = sizeof(T) == 4;
    const T X = x[1], Y = x[2], Z = x[3];
    const T R2v = X * X + Y * Y + Z * Z;
    const T w = R2v - s.a2;                           // R^2 - a^2
    T disc = tfma(T(4) * s.a2, Z * Z, w * w);         // (R^2 - a^2)^2 + 4 a^2 z^2
    disc = fmax(disc, T(1e-20));  // guards only the ring singularity, deep inside the horizon
    // FP32: square roots and their reciprocals from one hardware reciprocal square root each,
    // and 1 / rho, 1 / Dn from one reciprocal of their product (three transcendental
    // instructions instead of six; rho Dn ~ r^6 stays far inside the FP32 range for r < 1e6)
    T isq = T(0);
    const T sq = f32 ? disc * (isq = rsqrt_t(disc)) : sqrt(disc);
    T r2 = (w + sq) * T(0.5);
    r2 = fmax(r2, T(1e-10));
    T ir;
    const T r = f32 ? r2 * (ir = rsqrt_t(r2)) : sqrt(r2);
    if constexpr (!f32) ir = T(1) / r;
    const T rho = r2 + s.a2;
    const T Dn = tfma(s.a2 * Z, Z, r2 * r2);          // r^4 + a^2 z^2
    T irho, iDn;
    if constexpr (f32) {
        const T inv = rcp_t(rho * Dn);
        irho = Dn * inv;
        iDn = rho * inv;
    } else {
        irho = T(1) / rho;
        iDn = T(1) / Dn;
    }
    const T l1 = tfma(r, X, s.a * Y) * irho;          // (r x + a y) / (r^2 + a^2)
    const T l2 = tfma(r, Y, -s.a * X) * irho;         // (r y - a x) / (r^2 + a^2)
    const T l3 = Z * ir;                              // z / r
    const T q = T(2) * s.M * r - s.Q2;                // 2 M r - Q^2
    const T H = q * r2 * iDn;
    // k = (eta - H l l) p, l^mu = (-1, l1, l2, l3)
    T lp = -p[0] + l1 * p[1] + l2 * p[2] + l3 * p[3];
    T Hl = H * lp;
    k[0] = -p[0] + Hl;
    k[1] = p[1] - Hl * l1;
    k[2] = p[2] - Hl * l2;
    k[3] = p[3] - Hl * l3;
    T lam3 = T(0);
    if (s.Lambda != T(0)) {
        // g = g0 + c e0 e0^T with c = -Lambda R^2 / 3; g^-1 by Sherman-Morrison.
        T c = -s.Lambda * R2v / T(3);
        T u0 = -T(1) - H;
        T fac = c * k[0] / (T(1) + c * u0);
        k[0] -= fac * u0;
        k[1] -= fac * H * l1;
        k[2] -= fac * H * l2;
        k[3] -= fac * H * l3;
        lam3 = s.Lambda / T(3);
    }
    T lk = k[0] + l1 * k[1] + l2 * k[2] + l3 * k[3];
    T hl = T(0.5) * lk * lk, Hlk = H * lk, k00 = k[0] * k[0];
    // gradient of r
    const T rs = f32 ? r * isq : r / sq;
    const T drx = X * rs, dry = Y * rs, drz = f32 ? Z * rho * ir * isq : Z * rho * ir / sq;
    // H_r = (N_r - H Dn_r) / Dn with N = q r^2, N_r = r (6 M r - 2 Q^2), Dn_r = 4 r^3;
    // H_z = -2 a^2 z H / Dn
    const T Hr = (r * (T(6) * s.M * r - T(2) * s.Q2) - H * T(4) * r2 * r) * iDn;
    const T Hz = -T(2) * s.a2 * Z * H * iDn;
    const T A1 = (k[1] * r - k[2] * s.a) * irho, A2 = (k[1] * s.a + k[2] * r) * irho, A3 = k[3] * ir;
    const T B = (k[1] * X + k[2] * Y - T(2) * r * (k[1] * l1 + k[2] * l2)) * irho - k[3] * l3 * ir;
    const T Cr = hl * Hr + Hlk * B;
    dp[0] = Cr * drx + Hlk * A1 - lam3 * X * k00;
    dp[1] = Cr * dry + Hlk * A2 - lam3 * Y * k00;
    dp[2] = Cr * drz + hl * Hz + Hlk * A3 - lam3 * Z * k00;
    return p[0] * k[0] + p[1] * k[1] + p[2] * k[2] + p[3] * k[3];
}

// Ellis wormhole, orbital-plane reduction (t, l, psi, w): g = diag(-1, 1, l^2 + b0^2, 1).
template <class T>
SIRIUS_HD __forceinline__ T rhs_mt(const Spacetime<T>& s, const T x[4], const T p[4], T k[4], T dp[3]) {
    T r2 = x[1] * x[1] + s.b02;
    k[0] = -p[0];
    k[1] = p[1];
    k[2] = p[2] / r2;
    k[3] = p[3];
    dp[0] = x[1] * k[2] * k[2];
    dp[1] = T(0);
    dp[2] = T(0);
    return p[0] * k[0] + p[1] * k[1] + p[2] * k[2] + p[3] * k[3];
}

// Alcubierre, static bubble: g^{00} = -1, g^{01} = -v f, g^{11} = 1 - v^2 f^2.
template <class T>
SIRIUS_HD __forceinline__ T rhs_warp(const Spacetime<T>& s, const T x[4], const T p[4], T k[4], T dp[3]) {
    T R2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
    R2 = R2 < T(1e-20) ? T(1e-20) : R2;
    T rs = sqrt(R2);
    T dfdr;
    T f = warp_shape(s, rs, dfdr);
    T vf = s.wv * f;
    k[0] = -p[0] - vf * p[1];
    k[1] = -vf * p[0] + (T(1) - vf * vf) * p[1];
    k[2] = p[2];
    k[3] = p[3];
    T c = s.wv * dfdr / rs * (vf * k[0] * k[0] - k[0] * k[1]);
    dp[0] = c * x[1];
    dp[1] = c * x[2];
    dp[2] = c * x[3];
    return p[0] * k[0] + p[1] * k[1] + p[2] * k[2] + p[3] * k[3];
}

template <class T, int FAM>
SIRIUS_HD __forceinline__ T rhs(const Spacetime<T>& s, const T x[4], const T p[4], T k[4], T dp[3]) {
    if constexpr (FAM == 1) return rhs_mt<T>(s, x, p, k, dp);
    else if constexpr (FAM == 2) return rhs_warp<T>(s, x, p, k, dp);
    else return rhs_ks<T>(s, x, p, k, dp);
}

// k = g^{mu nu} p_nu from metric values only (no gradient): the FP64 null-constraint
// re-normalisation and drift report of the FP32 path use this lighter form.
template <class T, int FAM>
SIRIUS_HD __forceinline__ void raise_index(const Spacetime<T>& s, const T x[4], const T p[4], T k[4]) {
    if constexpr (FAM == 1) {
        k[0] = -p[0]; k[1] = p[1]; k[2] = p[2] / (x[1] * x[1] + s.b02); k[3] = p[3];
    } else if constexpr (FAM == 2) {
        T rs = sqrt(x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
        T dfdr;
        T vf = s.wv * warp_shape(s, rs, dfdr);
        k[0] = -p[0] - vf * p[1];
        k[1] = -vf * p[0] + (T(1) - vf * vf) * p[1];
        k[2] = p[2];
        k[3] = p[3];
    } else {
        T r = kerr_r(s, x[1], x[2], x[3]);
        T r2 = r * r, den = T(1) / (r2 + s.a2);
        T l1 = (r * x[1] + s.a * x[2]) * den, l2 = (r * x[2] - s.a * x[1]) * den, l3 = x[3] / r;
        T H = (T(2) * s.M * r - s.Q2) * r2 / (r2 * r2 + s.a2 * x[3] * x[3]);
        T Hl = H * (-p[0] + l1 * p[1] + l2 * p[2] + l3 * p[3]);
        k[0] = -p[0] + Hl;
        k[1] = p[1] - Hl * l1;
        k[2] = p[2] - Hl * l2;
        k[3] = p[3] - Hl * l3;
        if (s.Lambda != T(0)) {
            T c = -s.Lambda * (x[1] * x[1] + x[2] * x[2] + x[3] * x[3]) / T(3);
            T u0 = -T(1) - H;
            T fac = c * k[0] / (T(1) + c * u0);
            k[0] -= fac * u0;
            k[1] -= fac * H * l1;
            k[2] -= fac * H * l2;
            k[3] -= fac * H * l3;
        }
    }
}

// Derivative of the state: dx (= k), dp (spatial), and p.k.
template <class T>
struct Der {
    T k[4];
    T f[3];
    T pk;
};

template <class T, int FAM>
SIRIUS_HD __forceinline__ void eval(const Spacetime<T>& s, const T x[4], const T p[4], Der<T>& d) {
    d.pk = rhs<T, FAM>(s, x, p, d.k, d.f);
}

// Stage point y0 + inc, and the step update y0 + inc: with a compensated state (lo non-null) the
// low part lo[j] joins the increment, and the update's sum is error-free (Knuth's two-sum: hi + lo1[j] equals y0 + lo[j] + inc
// exactly, inc as rounded).
template <class T>
SIRIUS_HD __forceinline__ T stage_pt(T y0, T inc, const T* lo, int j) {
    return lo ? y0 + (lo[j] + inc) : y0 + inc;
}

template <class T>
SIRIUS_HD __forceinline__ T sum_lo(T y0, T inc, const T* lo, T* lo1, int j) {
    if (!lo) return y0 + inc;
    T t = lo[j] + inc;
    T hi = y0 + t;
    T bb = hi - y0;
    lo1[j] = (y0 - (hi - bb)) + (t - bb);
    return hi;
}

// ------------------------------------------------------------------------------------------
// Dormand-Prince 5(4), PHGD001A.cpp:16-79. p_t is a constant of motion and is not stepped.
// ------------------------------------------------------------------------------------------
// TIME = false skips the coordinate time t (x[0]), which no equation depends on: the FP32
// path integrates the 6 components (x, y, z, p_x, p_y, p_z) and measures its error on x, y, z.
// dense (optional, 6 values): the step's dense-output term r5 for x, y, z, p_x, p_y, p_z, with
// which dense_eval interpolates anywhere inside the step without new evaluations.
// lo (optional, 8 values: x^mu then p_mu): low parts of a compensated state, the state being
// (x0 + lo[0..3], p0 + lo[4..7]). They enter every stage point, and the step's increment is
// accumulated with an error-free sum whose low parts go to lo1. The FP32 path keeps its state
// this way, so rounding of the position and momentum does not accumulate along the ray.
template <class T, int FAM, bool TIME = true>
SIRIUS_HD __forceinline__ void dp45_step(const Spacetime<T>& s, const T x0[4], const T p0[4], const Der<T>& d1,
                                 T h, T x1[4], T p1[4], Der<T>& d7, T err[4], bool want_err,
                                 T* dense = nullptr, const T* lo = nullptr, T* lo1 = nullptr) {
    constexpr double a21 = 1.0 / 5.0;
    constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
    constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
    constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0, a54 = -212.0 / 729.0;
    constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0, a64 = 49.0 / 176.0,
                     a65 = -5103.0 / 18656.0;
    constexpr double b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0, b5 = -2187.0 / 6784.0,
                     b6 = 11.0 / 84.0;
    constexpr double e1 = 71.0 / 57600.0, e3 = -71.0 / 16695.0, e4 = 71.0 / 1920.0, e5 = -17253.0 / 339200.0,
                     e6 = 22.0 / 525.0, e7 = -1.0 / 40.0;
    Der<T> d2, d3, d4, d5, d6;
    T xt[4], pt[4];
    pt[0] = p0[0];
    constexpr int I0 = TIME ? 0 : 1;
    xt[0] = x0[0];
    x1[0] = x0[0];
#pragma unroll
    for (int i = I0; i < 4; ++i) xt[i] = stage_pt(x0[i], h * (T(a21) * d1.k[i]), lo, i);
#pragma unroll
    for (int i = 0; i < 3; ++i) pt[i + 1] = stage_pt(p0[i + 1], h * (T(a21) * d1.f[i]), lo, 5 + i);
    eval<T, FAM>(s, xt, pt, d2);
#pragma unroll
    for (int i = I0; i < 4; ++i) xt[i] = stage_pt(x0[i], h * (T(a31) * d1.k[i] + T(a32) * d2.k[i]), lo, i);
#pragma unroll
    for (int i = 0; i < 3; ++i) pt[i + 1] = stage_pt(p0[i + 1], h * (T(a31) * d1.f[i] + T(a32) * d2.f[i]), lo, 5 + i);
    eval<T, FAM>(s, xt, pt, d3);
#pragma unroll
    for (int i = I0; i < 4; ++i) xt[i] = stage_pt(x0[i], h * (T(a41) * d1.k[i] + T(a42) * d2.k[i] + T(a43) * d3.k[i]), lo, i);
#pragma unroll
    for (int i = 0; i < 3; ++i) pt[i + 1] = stage_pt(p0[i + 1], h * (T(a41) * d1.f[i] + T(a42) * d2.f[i] + T(a43) * d3.f[i]), lo, 5 + i);
    eval<T, FAM>(s, xt, pt, d4);
#pragma unroll
    for (int i = I0; i < 4; ++i)
        xt[i] = stage_pt(x0[i], h * (T(a51) * d1.k[i] + T(a52) * d2.k[i] + T(a53) * d3.k[i] + T(a54) * d4.k[i]), lo, i);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        pt[i + 1] = stage_pt(p0[i + 1], h * (T(a51) * d1.f[i] + T(a52) * d2.f[i] + T(a53) * d3.f[i] + T(a54) * d4.f[i]), lo, 5 + i);
    eval<T, FAM>(s, xt, pt, d5);
#pragma unroll
    for (int i = I0; i < 4; ++i)
        xt[i] = stage_pt(x0[i], h * (T(a61) * d1.k[i] + T(a62) * d2.k[i] + T(a63) * d3.k[i] + T(a64) * d4.k[i] + T(a65) * d5.k[i]), lo, i);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        pt[i + 1] = stage_pt(p0[i + 1], h * (T(a61) * d1.f[i] + T(a62) * d2.f[i] + T(a63) * d3.f[i] + T(a64) * d4.f[i] + T(a65) * d5.f[i]), lo, 5 + i);
    eval<T, FAM>(s, xt, pt, d6);
    p1[0] = p0[0];
#pragma unroll
    for (int i = I0; i < 4; ++i)
        x1[i] = sum_lo(x0[i], h * (T(b1) * d1.k[i] + T(b3) * d3.k[i] + T(b4) * d4.k[i] + T(b5) * d5.k[i] +
                                   T(b6) * d6.k[i]), lo, lo1, i);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        p1[i + 1] = sum_lo(p0[i + 1], h * (T(b1) * d1.f[i] + T(b3) * d3.f[i] + T(b4) * d4.f[i] + T(b5) * d5.f[i] +
                                           T(b6) * d6.f[i]), lo, lo1, 5 + i);
    eval<T, FAM>(s, x1, p1, d7);
    if (want_err) {
        err[0] = T(0);
#pragma unroll
        for (int i = I0; i < 4; ++i)
            err[i] = h * (T(e1) * d1.k[i] + T(e3) * d3.k[i] + T(e4) * d4.k[i] + T(e5) * d5.k[i] + T(e6) * d6.k[i] +
                          T(e7) * d7.k[i]);
    }
    if (dense) {
        // Dormand-Prince continuous extension of order 4 (Hairer, Norsett & Wanner, Solving
        // Ordinary Differential Equations I, sec. II.6; their DOPRI5 code's dense output)
        constexpr double q1 = -12715105075.0 / 11282082432.0, q3 = 87487479700.0 / 32700410799.0,
                         q4 = -10690763975.0 / 1880347072.0, q5 = 701980252875.0 / 199316789632.0,
                         q6 = -1453857185.0 / 822651844.0, q7 = 69997945.0 / 29380423.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            dense[i] = h * (T(q1) * d1.k[i + 1] + T(q3) * d3.k[i + 1] + T(q4) * d4.k[i + 1] + T(q5) * d5.k[i + 1] +
                            T(q6) * d6.k[i + 1] + T(q7) * d7.k[i + 1]);
            dense[3 + i] = h * (T(q1) * d1.f[i] + T(q3) * d3.f[i] + T(q4) * d4.f[i] + T(q5) * d5.f[i] +
                                T(q6) * d6.f[i] + T(q7) * d7.f[i]);
        }
    }
}

// One component of the dense output at th in [0, 1] from its values y0, y1 at the step's ends,
// the end slopes k0, k1, the step length h and the step's r5 term:
//   y(th) = y0 + th (d + (1 - th) (b + th (c + (1 - th) r5))),
//   d = y1 - y0, b = hk0 - d, c = d - hk1 - b.
// It matches y and its slope at both ends; with r5 = 0 it is the cubic Hermite interpolant.
// Every multiply-add is an explicit fma, so the rounding does not depend on how the compiler
// contracts (and the same ray gives the same bits in every kernel that traces it).
template <class T>
SIRIUS_HD __forceinline__ T dense_eval(T y0, T y1, T h, T k0, T k1, T r5, T th) {
    T d = y1 - y0;
    T b = tfma(h, k0, -d);
    T c = d - tfma(h, k1, b);
    T t1 = T(1) - th;
    T u = tfma(t1, r5, c);
    u = tfma(th, u, b);
    u = tfma(t1, u, d);
    return tfma(th, u, y0);
}

// x^e for x > 0. FP32 on the device: the hardware base-2 logarithm and exponential (step-size
// factors are heuristics; a general powf spends ~40 instructions on range and special cases).
// FP64, and host builds of this header: pow.
template <class T>
SIRIUS_HD __forceinline__ T pow_pos(T x, T e) {
#ifdef __HIP_DEVICE_COMPILE__
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
#endif
    return pow(x, e);
}

// computeOptimalStep (PHGD001A.cpp:471-483): h * clamp(0.9 err^-1/5, 0.1, 2), clamped.
template <class T>
SIRIUS_HD __forceinline__ T next_step(T h, T err, T hmin, T hmax) {
    T hn;
    if (err < T(1e-15)) hn = h * T(2);
    else {
        T fac = T(0.9) * pow_pos(err, T(-0.2));
        fac = fac < T(0.1) ? T(0.1) : (fac > T(2) ? T(2) : fac);
        hn = h * fac;
    }
    hn = hn < hmin ? hmin : hn;
    hn = hn > hmax ? hmax : hn;
    return hn;
}

// Cubic Hermite position over a step, th in [0, 1].
template <class T>
SIRIUS_HD __forceinline__ void hermite(const T x0[4], const T k0[4], const T x1[4], const T k1[4], T h, T th,
                                        T out[4]) {
    T t2 = th * th, t3 = t2 * th;
    T h00 = T(2) * t3 - T(3) * t2 + T(1), h10 = t3 - T(2) * t2 + th, h01 = T(3) * t2 - T(2) * t3, h11 = t3 - t2;
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = h00 * x0[i] + h10 * h * k0[i] + h01 * x1[i] + h11 * h * k1[i];
}

enum { EV_Z = 0, EV_R = 1 };

template <class T>
SIRIUS_HD __forceinline__ T event_value(const Spacetime<T>& s, int kind, const T x[4], T target) {
    if (kind == EV_Z) return x[3];
    return event_radius(s, x) - target;
}

template <class T>
SIRIUS_HD __forceinline__ T locate_event(const Spacetime<T>& s, int kind, T target, const T x0[4], const T k0[4],
                                 const T x1[4], const T k1[4], T h) {
    constexpr int iters = sizeof(T) == 8 ? 52 : 24;
    T lo = T(0), hi = T(1), xm[4];
    bool neg_lo = event_value(s, kind, x0, target) < T(0);
    for (int it = 0; it < iters; ++it) {
        T mid = T(0.5) * (lo + hi);
        hermite(x0, k0, x1, k1, h, mid, xm);
        bool neg = event_value(s, kind, xm, target) < T(0);
        if (neg == neg_lo) lo = mid; else hi = mid;
    }
    return T(0.5) * (lo + hi);
}

// Position on the dense output of a step (x^0 is carried unchanged: the FP32 path does not
// integrate t).
template <class T>
SIRIUS_HD __forceinline__ void dense_pos(const T x0[4], const T k0[4], const T x1[4], const T k1[4], const T dn[6],
                                         T h, T th, T out[4]) {
    out[0] = x0[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) out[i] = dense_eval(x0[i], x1[i], h, k0[i], k1[i], dn[i - 1], th);
}

// Slope dy/dth of dense_eval's polynomial
//   y(th) = y0 + th d + th (1 - th) b + th^2 (1 - th) c + th^2 (1 - th)^2 r5.
template <class T>
SIRIUS_HD __forceinline__ T dense_slope(T y0, T y1, T h, T k0, T k1, T r5, T th) {
    T d = y1 - y0;
    T b = tfma(h, k0, -d);
    T c = d - tfma(h, k1, b);
    T t1 = T(1) - th;
    T u = T(2) * th * t1 * (t1 - th) * r5;
    u = tfma(th * (T(2) - T(3) * th), c, u);
    u = tfma(t1 - th, b, u);
    return d + u;
}

// Event location on the dense output by safeguarded Newton iteration: the root of the event
// function over th in [0, 1], where it changes sign. Each iterate stays inside the bracket that
// the signs seen so far keep (a step leaving it is replaced by the bracket's midpoint), starting
// from the linear interpolant's root. The bracket is closed: a converged iterate has just become
// one of its ends, and its Newton update (itself, to rounding) is kept, not bisected away. The
// iteration ends when a Newton step moves th by less than tol, or after 24 rounds (each of them
// at least as good as a halving: a step falling back to the midpoint halves the bracket, one
// staying inside converges quadratically). The event functions are smooth across a step (the
// plane z = 0, or a sphere of the event radius), so most crossings converge in two or three
// rounds, for a fraction of the instructions of 24 halvings: a wave runs this whenever one of its
// lanes has an event in the step, for as many rounds as its slowest lane needs.
template <class T>
SIRIUS_HD __forceinline__ T locate_event_newton(const Spacetime<T>& s, int kind, T target, const T x0[4],
                                                const T k0[4], const T x1[4], const T k1[4], const T dn[6], T h,
                                                T f0, T f1) {
    constexpr int max_iters = 24;
    const T tol = T(sizeof(T) == 4 ? 2e-7 : 1e-15);
    T lo = T(0), hi = T(1);
    const bool neg_lo = f0 < T(0);
    T den = f0 - f1;
    T th = den != T(0) ? f0 / den : T(0.5);
    th = th > T(0) && th < T(1) ? th : T(0.5);
    for (int it = 0; it < max_iters; ++it) {
        T f, df;
        if (kind == EV_Z) {
            f = dense_eval(x0[3], x1[3], h, k0[3], k1[3], dn[2], th);
            df = dense_slope(x0[3], x1[3], h, k0[3], k1[3], dn[2], th);
        } else {
            T xm[4], v[4];
            dense_pos(x0, k0, x1, k1, dn, h, th, xm);
#pragma unroll
            for (int i = 1; i < 4; ++i) v[i] = dense_slope(x0[i], x1[i], h, k0[i], k1[i], dn[i - 1], th);
            T rate;
            f = event_radius_rate(s, xm, v, rate) - target;
            df = rate;
        }
        if (f == T(0)) break;
        if ((f < T(0)) == neg_lo) lo = th; else hi = th;
        const T nt = df != T(0) ? th - f / df : T(-1);
        if (nt >= lo && nt <= hi) {
            const bool done = fabs(nt - th) <= tol;
            th = nt;
            if (done) break;
        } else {
            th = T(0.5) * (lo + hi);
        }
    }
    return th;
}

// Event location on the dense output: bisection of the event function over th in [0, 1]
// (the plane z = 0 needs only the z component).
template <class T>
SIRIUS_HD __forceinline__ T locate_event_dense(const Spacetime<T>& s, int kind, T target, const T x0[4],
                                               const T k0[4], const T x1[4], const T k1[4], const T dn[6], T h) {
    constexpr int iters = sizeof(T) == 8 ? 52 : 24;
    T lo = T(0), hi = T(1), xm[4];
    bool neg_lo = event_value(s, kind, x0, target) < T(0);
    for (int it = 0; it < iters; ++it) {
        T mid = T(0.5) * (lo + hi);
        T v;
        if (kind == EV_Z) v = dense_eval(x0[3], x1[3], h, k0[3], k1[3], dn[2], mid);
        else {
            dense_pos(x0, k0, x1, k1, dn, h, mid, xm);
            v = event_value(s, kind, xm, target);
        }
        bool neg = v < T(0);
        if (neg == neg_lo) lo = mid; else hi = mid;
    }
    return T(0.5) * (lo + hi);
}

// ------------------------------------------------------------------------------------------
// Camera and initial state (CMBS001A.h:120-185, GTRC001A.cpp:15-106)
// ------------------------------------------------------------------------------------------
// Direction through the image point (u, v) in pixel units (a pixel's centre is (px + 1/2, py + 1/2)).
template <class T>
SIRIUS_HD __forceinline__ void camera_dir_uv(const CamConst& c, T u, T v, T nsph[3]) {
    T sx = (T(2) * u / T(c.width) - T(1)) * T(c.aspect);
    T sy = T(1) - T(2) * v / T(c.height);
    T dx = sx * T(c.tan_half), dy = sy * T(c.tan_half), dz = T(-1);
    T cr = T(c.croll), sr = T(c.sroll), cp = T(c.cpitch), sp = T(c.spitch), cy = T(c.cyaw), sy_ = T(c.syaw);
    T rx = dx * cr - dy * sr, ry = dx * sr + dy * cr;
    T ry2 = ry * cp - dz * sp, rz = ry * sp + dz * cp;
    T rx2 = rx * cy + rz * sy_, rz2 = -rx * sy_ + rz * cy;
    T il = T(1) / sqrt(rx2 * rx2 + ry2 * ry2 + rz2 * rz2);
    nsph[0] = rz2 * il;
    nsph[1] = -ry2 * il;
    nsph[2] = rx2 * il;
}

template <class T>
SIRIUS_HD __forceinline__ void camera_dir(const CamConst& c, int px, int py, T nsph[3]) {
    camera_dir_uv<T>(c, T(px) + T(0.5), T(py) + T(0.5), nsph);
}

// Future-directed k^0 of MTTN001A.cpp:171-237 (same root choice and fallbacks). With
// near > 0 (re-normalisation of a ray in flight) the root nearest the current k^0 is taken
// instead: inside the ergoregion g_00 > 0 and both roots can be positive, where the
// reference's first-positive-root rule can jump to the wrong branch.
template <class T>
SIRIUS_HD __forceinline__ T null_k0(const T g[4][4], const T k[4], T near = T(0)) {
    T A = g[0][0], B = T(0), C = T(0);
    for (int i = 1; i < 4; ++i) B += T(2) * g[0][i] * k[i];
    for (int i = 1; i < 4; ++i)
        for (int j = 1; j < 4; ++j) C += g[i][j] * k[i] * k[j];
    if (fabs(A) < T(1e-15)) return fabs(B) > T(1e-15) ? -C / B : T(1);
    T disc = B * B - T(4) * A * C;
    if (disc < T(0)) return T(1);
    T sq = sqrt(disc);
    T kp = (-B + sq) / (T(2) * A), km = (-B - sq) / (T(2) * A);
    if (near > T(0)) return fabs(kp - near) < fabs(km - near) ? kp : km;
    if (kp > T(0)) return kp;
    if (km > T(0)) return km;
    return kp > km ? kp : km;
}

// Wormhole orbital-plane frame of a ray: n_r (radial) and n_t (tangential magnitude).
template <class T>
SIRIUS_HD __forceinline__ void mt_radial_tangential(const CamConst& c, const T ns[3], T& nr, T& nt) {
    nr = ns[0];
    nt = sqrt(ns[1] * ns[1] + ns[2] * ns[2]);
}

// Unit initial direction of pixel (px, py) in the Cartesian frame (Kerr-Schild and warp cameras).
template <class T>
SIRIUS_HD __forceinline__ void camera_cart_dir(const CamConst& c, int px, int py, T n[3]) {
    T ns[3];
    camera_dir<T>(c, px, py, ns);
#pragma unroll
    for (int i = 0; i < 3; ++i) n[i] = ns[0] * T(c.er[i]) + ns[1] * T(c.et[i]) + ns[2] * T(c.ep[i]);
}

template <class T>
SIRIUS_HD __forceinline__ void init_state_ns(const CamConst& c, const Spacetime<T>& s, const T ns[3], T x[4], T p[4]) {
    T k[4], g[4][4];
    if (c.family == 1) {
        T nr, nt;
        mt_radial_tangential(c, ns, nr, nt);
        T r = T(c.r_obs);
        x[0] = T(0); x[1] = T(c.l_obs); x[2] = T(0); x[3] = T(0);
        k[0] = T(1); k[1] = nr; k[2] = nt / r; k[3] = T(0);
    } else {
        x[0] = T(0);
        k[0] = T(0);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x[i + 1] = T(c.r_obs * c.er[i]);
            k[i + 1] = ns[0] * T(c.er[i]) + ns[1] * T(c.et[i]) + ns[2] * T(c.ep[i]);  // unit Cartesian direction
        }
    }
    // the metric at the observer is the camera's: computed once on the host in FP64
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) g[mu][nu] = T(c.g_obs[4 * mu + nu]);
    if (c.family != 1) k[0] = null_k0(g, k);
    for (int mu = 0; mu < 4; ++mu) {
        T acc = T(0);
        for (int nu = 0; nu < 4; ++nu) acc += g[mu][nu] * k[nu];
        p[mu] = acc;
    }
}

template <class T>
SIRIUS_HD __forceinline__ void init_state(const CamConst& c, const Spacetime<T>& s, int px, int py, T x[4], T p[4]) {
    T ns[3];
    camera_dir<T>(c, px, py, ns);
    init_state_ns<T>(c, s, ns, x, p);
}

// Initial state of the ray through the image point (u, v) (pixel units).
template <class T>
SIRIUS_HD __forceinline__ void init_state_uv(const CamConst& c, const Spacetime<T>& s, T u, T v, T x[4], T p[4]) {
    T ns[3];
    camera_dir_uv<T>(c, u, v, ns);
    init_state_ns<T>(c, s, ns, x, p);
}

// ------------------------------------------------------------------------------------------
// Emission (evaluated in FP64 on both paths; once per ray or per volume node)
// ------------------------------------------------------------------------------------------
// Page & Thorne (1974) closed-form f(r) with the camera's precomputed roots (M = 1).
SIRIUS_HD __forceinline__ double page_thorne_f(const CamConst& c, double r) {
    double x = sqrt(r), x0 = c.pt_x0;
    if (x <= x0) return 0.0;
    double br = x - x0 - 1.5 * c.a * log(x / x0);
#pragma unroll
    for (int i = 0; i < 3; ++i) br -= c.pt_c[i] * log((x - c.pt_x[i]) / (x0 - c.pt_x[i]));
    return 1.5 / (x * x * (x * x * x - 3.0 * x + 2.0 * c.a)) * br;
}

// Kepler angular velocity of a prograde circular Kerr-Newman orbit at BL radius r.
template <class T = double>
SIRIUS_HD __forceinline__ T kepler_omega(const CamConst& c, T r) {
    T sq = sqrt(fmax(T(c.M) * r - T(c.Q2), T(0)));
    return sq / (r * r + T(c.a) * sq);
}

// u^t of the circular emitter at x with angular velocity Om (u = u^t (d_t + Om d_phi)).
template <class T = double>
SIRIUS_HD __forceinline__ T emitter_ut(const CamConst& c, const Spacetime<T>& s, const T x[4], T Om) {
    T g[4][4];
    metric_cov(s, x, g);
    T v[4] = {T(1), -Om * x[2], Om * x[1], T(0)};
    T vv = T(0);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) vv += g[i][j] * v[i] * v[j];
    return vv < T(0) ? T(1) / sqrt(-vv) : T(-1);
}

// Spectral bin centres x_k = x_min (x_max / x_min)^((k + 1/2) / n_bins): the first centre and the
// ratio of neighbours (each centre is the previous one times the ratio). The host computes both
// once per batch (bin_centres); kernels read them from the parameters instead of evaluating an
// FP64 exp and log per ray or per volumetric step.
inline void bin_centres(double x_min, double x_max, int n_bins, double& x0, double& q) {
    if (n_bins <= 0) { q = 1.0; x0 = x_min; return; }
    q = std::exp(std::log(x_max / x_min) / n_bins);
    x0 = x_min * std::sqrt(q);
}
SIRIUS_HD __forceinline__ void bin_geometry(const TraceParams& tp, double& x0, double& q) {
    q = tp.bin_q;
    x0 = tp.bin_x0;
}

// Thin disk: g = (p.u_obs)/(p.u_emit), I_k = x_k^3 / (e^{x_k/(g tau)} - 1). The redshift and the
// temperature are FP64; the Planck bins are FP64, or FP32 (bins32: rays of the FP32 path, whose
// intensities are held to a 1e-3 relative tolerance).
SIRIUS_HD __forceinline__ void disk_emission(const CamConst& c, const TraceParams& tp, const double x[4], const double p[4],
                                     double r, double& gf, float* I, int stride, bool bins32) {
    Spacetime<double> s = make_spacetime<double>(c);
    double Om = kepler_omega(c, r);
    double ut = emitter_ut(c, s, x, Om);
    double pphi = x[1] * p[2] - x[2] * p[1];
    gf = (p[0] * c.obs_norm) / (ut * (p[0] + Om * pphi));
    double f = page_thorne_f(c, r) / r;
    double tau = (c.fmax > 0 && f > 0) ? sqrt(sqrt(f / c.fmax)) : 0.0;  // (f / f_max)^(1/4)
    const bool live = tau > 0 && gf > 0 && ut > 0;
    double xk, q;
    bin_geometry(tp, xk, q);
    const double inv = live ? 1.0 / (gf * tau) : 0.0;
    for (int k = 0; k < tp.n_bins; ++k, xk *= q) {
        float v = 0.f;
        if (live) {
            if (bins32) {
                float xf = (float)xk;
                v = xf * xf * xf / expm1f((float)(xk * inv));
            } else {
                v = (float)(xk * xk * xk / expm1(xk * inv));
            }
        }
        I[(int64_t)k * stride] = v;
    }
}

// ------------------------------------------------------------------------------------------
// Critical-curve distance. Per camera the host builds a radial table of the curve in the
// (alpha, beta) plane: tab[0] = alpha0 (the centre), tab[1 + j] = rho_c(pi j / (N - 1)).
// alpha = -L/(E sin th), beta = p_th/E at the ray's start; distance = rho - rho_c(psi).
// ------------------------------------------------------------------------------------------
#define SIRIUS_CC_TABLE 512

// Impact parameters of a ray on the observer's sky (Bardeen): alpha = -L_z / (E sin theta),
// beta = p_theta / E, from the conserved E = -p_t and L_z = x p_y - y p_x and the polar
// momentum at the ray's position.
template <class T>
SIRIUS_HD __forceinline__ void sky_alpha_beta(const Spacetime<T>& s, const T x[4], const T p[4], T& alpha, T& beta) {
    T E = -p[0];
    T L = x[1] * p[2] - x[2] * p[1];
    T r = kerr_r(s, x[1], x[2], x[3]);
    T ct = x[3] / r, st = sqrt(fmax(T(1) - ct * ct, T(1e-30)));
    T pth = (ct / st) * (x[1] * p[1] + x[2] * p[2]) - r * st * p[3];
    alpha = -L / (E * st);
    beta = pth / E;
}

template <class T>
SIRIUS_HD __forceinline__ T crit_distance(const float* tab, const Spacetime<T>& s, const T x[4], const T p[4]) {
    T alpha, beta;
    sky_alpha_beta(s, x, p, alpha, beta);
    T da = alpha - T(tab[0]), b = fabs(beta);
    T psi = atan2(b, da), rho = sqrt(da * da + b * b);
    T u = psi / T(3.141592653589793) * T(SIRIUS_CC_TABLE - 1);
    int j = (int)u;
    j = j < 0 ? 0 : (j > SIRIUS_CC_TABLE - 2 ? SIRIUS_CC_TABLE - 2 : j);
    T w = u - T(j);
    return rho - (T(tab[1 + j]) + w * (T(tab[2 + j]) - T(tab[1 + j])));
}

}  // namespace sirius
