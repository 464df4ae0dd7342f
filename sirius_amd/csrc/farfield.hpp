// farfield.hpp — the analytic reference of the far-field surrogate: the exact null geodesic of the
// Schwarzschild spacetime (M = 1) through a ray's initial position and coordinate direction.
//
// A Schwarzschild ray stays in the plane of its initial position and direction. With u = 1/r and
// b = L / E its orbit obeys (du/dphi)^2 = P(u) = 1/b^2 - u^2 + 2 u^3 = 2 (u - u1)(u - u2)(u - u3),
// u1 < 0 < u2 < u3 for b above the critical sqrt(27): an inbound ray sweeps
//   I(u_a) = int_{u_a}^{u2} du / sqrt(P(u))
// from radius 1/u_a to its turning point 1/u2 and the same from there out. With u = u2 - (u2 - u_a) t^2
// the integrand loses its inverse square root at the turning point, and a 16-point Gauss-Legendre
// rule reaches the working precision away from the critical curve. Kerr-Schild coordinates of
// a = 0 are the Schwarzschild ones in space, so this is the integrator's own answer at a = 0: the
// surrogate learns only what spin, charge and the cosmological term add to it. The same orbit
// decides which rays the surrogate may answer: those whose crossings of the equatorial plane all
// lie outside the thin disk (with a margin for what the spin moves).
#pragma once

#include "geodesic.hpp"

namespace sirius {

// Gauss-Legendre nodes and weights on [0, 1], 8 points (the equatorial-crossing test, which needs
// radii to ~1e-4 M) and 16 points (the exit direction).
constexpr double kFarGlT[16] = {0.005299532504175031, 0.0277124884633837,  0.06718439880608412, 0.1222977958224985,
                                0.19106187779867811,  0.2709916111713863,  0.35919822461037054, 0.4524937450811813,
                                0.5475062549188188,   0.6408017753896295,  0.7290083888286136,  0.8089381222013219,
                                0.8777022041775016,   0.9328156011939159,  0.9722875115366163,  0.994700467495825};
constexpr double kFarGlW[16] = {0.013576229705877019, 0.031126761969323853, 0.047579255841246296, 0.062314485627767015,
                                0.07479799440828838,  0.08457825969750131,  0.0913017075224618,   0.09472530522753429,
                                0.09472530522753429,  0.0913017075224618,   0.08457825969750131,  0.07479799440828838,
                                0.062314485627767015, 0.047579255841246296, 0.031126761969323853, 0.013576229705877019};

constexpr double kFarGl8T[8] = {0.019855071751231912, 0.10166676129318664, 0.2372337950418355, 0.4082826787521751,
                                0.5917173212478248,   0.7627662049581645,  0.8983332387068134, 0.9801449282487681};
constexpr double kFarGl8W[8] = {0.050614268145188344, 0.11119051722668717, 0.15685332293894352, 0.18134189168918088,
                                0.18134189168918088,  0.15685332293894352, 0.11119051722668717, 0.050614268145188344};

template <class T>
struct FarOrbit {
    T rh[3], et[3];     // orbit plane: unit initial position, unit tangential direction of travel
    T b;                // impact parameter L / E
    T u1, u2, u3;       // roots of P (u2: the turning point 1 / r_min)
    T u0, ue;           // 1 / r at the start and at the exit sphere
    T phi_in, phi_out;  // sweep from the start to the turning point, and from it to the exit sphere
    bool ok;            // scattered (b above the critical value) and inbound at the start
};

// 1 / sqrt(2 (u - u1)(u3 - u)): the smooth part of the orbit integrand.
template <class T>
SIRIUS_HD __forceinline__ T far_g(const FarOrbit<T>& o, T u) {
    T q = T(2) * (u - o.u1) * (o.u3 - u);
    return rsqrt_t(q > T(1e-30) ? q : T(1e-30));
}

// I(u2 - sigma^2) = 2 sigma int_0^1 g(u2 - sigma^2 t^2) dt, by the NQ-point rule (8 or 16).
template <int NQ, class T>
SIRIUS_HD __forceinline__ T far_sweep_sigma(const FarOrbit<T>& o, T sigma) {
    T acc = T(0);
    const T s2 = sigma * sigma;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const T t = T(NQ == 8 ? kFarGl8T[k] : kFarGlT[k]);
        acc = tfma(T(NQ == 8 ? kFarGl8W[k] : kFarGlW[k]), far_g(o, o.u2 - s2 * t * t), acc);
    }
    return T(2) * sigma * acc;
}

template <int NQ, class T>
SIRIUS_HD __forceinline__ T far_sweep(const FarOrbit<T>& o, T ua) {
    T d = o.u2 - ua;
    return far_sweep_sigma<NQ, T>(o, sqrt(d > T(0) ? d : T(0)));
}

// The orbit of the ray starting at x0 (Cartesian, |x0| = r0) with unit coordinate direction n;
// its sweeps by the NQ-point rule.
template <int NQ, class T>
SIRIUS_HD __forceinline__ FarOrbit<T> far_orbit(const T x0[3], const T n[3], T r_exit) {
    FarOrbit<T> o;
    const T r0 = sqrt(x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2]);
    T nr = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        o.rh[i] = x0[i] / r0;
        nr += n[i] * o.rh[i];
    }
    T nt2 = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        o.et[i] = n[i] - nr * o.rh[i];
        nt2 += o.et[i] * o.et[i];
    }
    const T nt = sqrt(nt2);
#pragma unroll
    for (int i = 0; i < 3; ++i) o.et[i] = nt > T(0) ? o.et[i] / nt : T(0);
    // E^2 = rdot^2 + (1 - 2/r) (r phidot)^2 and L = r (r phidot), with (rdot, r phidot) along (n_r, n_t)
    o.b = r0 * nt / sqrt(nr * nr + (T(1) - T(2) / r0) * nt2);
    const T c = T(1) / (o.b * o.b);
    o.ok = nr < T(0) && c < T(1.0 / 27.0) && nt > T(0);
    // roots of u^3 - u^2 / 2 + c / 2 by the trigonometric form (u = t + 1/6)
    T arg = T(1) - T(54) * c;
    arg = arg < T(-1) ? T(-1) : (arg > T(1) ? T(1) : arg);
    const T th = acos(arg) / T(3);
    o.u3 = cos(th) / T(3) + T(1.0 / 6.0);
    o.u2 = cos(th - T(2.0943951023931957)) / T(3) + T(1.0 / 6.0);
    o.u1 = cos(th - T(4.1887902047863914)) / T(3) + T(1.0 / 6.0);
    o.u0 = T(1) / r0;
    o.ue = T(1) / r_exit;
    o.phi_in = o.ok ? far_sweep<NQ, T>(o, o.u0) : T(0);
    o.phi_out = o.ok ? far_sweep<NQ, T>(o, o.ue) : T(0);
    return o;
}

// Unit coordinate direction at the exit sphere (the integrator's escape direction at a = 0).
template <class T>
SIRIUS_HD __forceinline__ void far_exit_dir(const FarOrbit<T>& o, T d[3]) {
    const T Phi = o.phi_in + o.phi_out;
    const T cP = cos(Phi), sP = sin(Phi);
    const T vr2 = T(1) - o.b * o.b * o.ue * o.ue * (T(1) - T(2) * o.ue);
    const T vr = sqrt(vr2 > T(0) ? vr2 : T(0)), vt = o.b * o.ue;
    T nn = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const T re = cP * o.rh[i] + sP * o.et[i], pe = -sP * o.rh[i] + cP * o.et[i];
        d[i] = vr * re + vt * pe;
        nn += d[i] * d[i];
    }
    const T inv = rsqrt_t(nn);
#pragma unroll
    for (int i = 0; i < 3; ++i) d[i] *= inv;
}

// 1 / r where the orbit's in-plane azimuth (from the start) is phi, 0 <= phi <= phi_in + phi_out:
// Newton's method on sigma, u = u2 - sigma^2, for I = phi_in - phi (inbound) or phi - phi_in
// (outbound); dI/dsigma = 2 g(u).
template <int NQ, class T>
SIRIUS_HD __forceinline__ T far_u_at(const FarOrbit<T>& o, T phi) {
    const bool inbound = phi <= o.phi_in;
    const T target = inbound ? o.phi_in - phi : phi - o.phi_in;
    const T ua = inbound ? o.u0 : o.ue;
    T smax2 = o.u2 - ua;
    const T smax = sqrt(smax2 > T(0) ? smax2 : T(0));
    T sigma = target / (T(2) * far_g(o, o.u2));
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        sigma = sigma < T(0) ? T(0) : (sigma > smax ? smax : sigma);
        const T f = far_sweep_sigma<NQ, T>(o, sigma) - target;
        sigma -= f / (T(2) * far_g(o, o.u2 - sigma * sigma));
    }
    sigma = sigma < T(0) ? T(0) : (sigma > smax ? smax : sigma);
    return o.u2 - sigma * sigma;
}

// Whether any crossing of the equatorial plane along the orbit (start to exit sphere) lies at a
// radius within [r_lo, r_hi]: the plane meets the orbit's plane along its line of nodes, at
// in-plane azimuths phi_n + k pi.
template <int NQ, class T>
SIRIUS_HD __forceinline__ bool far_crosses_annulus(const FarOrbit<T>& o, T r_lo, T r_hi) {
    const T rz = o.rh[2], ez = o.et[2];
    if (fabs(rz) < T(1e-6) && fabs(ez) < T(1e-6)) return true;  // the orbit lies in the disk's plane
    T phin = atan2(-rz, ez);
    if (phin < T(0)) phin += T(3.141592653589793);
    const T Phi = o.phi_in + o.phi_out;
    for (int k = 0; k < 3; ++k) {
        const T phi = phin + T(k) * T(3.141592653589793);
        if (phi >= Phi) break;
        const T r = T(1) / far_u_at<NQ, T>(o, phi);
        if (r >= r_lo && r <= r_hi) return true;
    }
    return false;
}

}  // namespace sirius
