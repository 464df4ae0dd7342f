// bundle_kernels.hip — ray bundles: each pixel's ray with its geodesic deviation.
//
// The reference follows two deviation vectors xi_1, xi_2 (seeded along the camera's right and
// up axes) beside the central ray with D^2 xi / dlam^2 = R^mu_{nu rho sigma} k^nu xi^rho k^sigma,
// either with a tidal approximation or a finite-difference Riemann tensor
// (RDOP002A.cu:4239-4354, 5014-5047, 5364-5439), and reads the beam's ellipse off their Gram
// matrix. Here the deviation is the linearisation of the ray's own flow: in Hamiltonian form
// the geodesic deviation equation is the tangent (variational) equation of
// (dx/dlam, dp/dlam) = (dH/dp, -dH/dx), so the two deviation vectors are the derivatives of
// the state (x^mu, p_mu) with respect to the image coordinates (u, v) of the ray's pixel. They
// are carried by dual numbers: the camera, the initial state and every DP45 step run on
// Dual<double, 1> values seeded with d/du, and again with d/dv, so each accepted step advances the
// deviation by the exact derivative of that step's map (its length held; the length is chosen
// on the central ray alone). At the end the derivative of the outcome follows: of the escape
// direction (theta, phi) for an escaped ray (rad per pixel), or of the equatorial hit point
// (X, Y) for a ray ending on the thin disk (M per pixel), the deviation there projected along
// the ray onto the plane.
//
// One thread per ray, FP64 values and tangents.
#include <hip/hip_runtime.h>

#include "../../include/sirius_hip.h"
#include "geodesic.hpp"
#include "sirius_bundle.hpp"

namespace sirius {

namespace {

SIRIUS_HD __forceinline__ int find_cam(const CamConst* cams, int n_cam, int64_t id) {
    int c = 0;
    while (c + 1 < n_cam && cams[c + 1].offset <= id) ++c;
    return c;
}

// Escape angles of the final state (x, p) of the ray through the image point (u, v), as the
// trace's outputs define them (the wormhole's orbital plane depends on the pixel).
template <int FAM>
SIRIUS_HD __forceinline__ void escape_angles(const CamConst& c, const Spacetime<double>& s, double u, double v,
                                             const double x[4], const double p[4], double& th, double& ph) {
    double k[4];
    raise_index<double, FAM>(s, x, p, k);
    double dir[3];
    if constexpr (FAM == 1) {
        double ns[3];
        camera_dir_uv<double>(c, u, v, ns);
        double nt = sqrt(ns[1] * ns[1] + ns[2] * ns[2]);
        double e2[3];
        for (int i = 0; i < 3; ++i) e2[i] = nt > 1e-300 ? (ns[1] * c.et[i] + ns[2] * c.ep[i]) / nt : c.et[i];
        double l = x[1], psi = x[2], r = sqrt(l * l + c.b0 * c.b0);
        double drdl = (l / r) * k[1], rdpsi = p[2] / r;
        double cp = cos(psi), sp = sin(psi);
        for (int i = 0; i < 3; ++i) {
            double e = cp * c.er[i] + sp * e2[i], t = -sp * c.er[i] + cp * e2[i];
            dir[i] = drdl * e + rdpsi * t;
        }
    } else {
        dir[0] = k[1];
        dir[1] = k[2];
        dir[2] = k[3];
    }
    double n = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    th = acos(fmin(1.0, fmax(-1.0, dir[2] / n)));
    ph = atan2(dir[1], dir[0]);
}

// Difference scale for the escape direction along a deviation t at state y: a perturbation of
// 1e-4 pixel, or less where the deviation has grown, so that eps |t| stays below 1e-6 of the
// state's scale.
SIRIUS_HD __forceinline__ double fd_eps(const double x[4], const double p[4], const double t[8]) {
    double ys = 1.0, tn = 0.0;
    for (int i = 1; i < 4; ++i) ys = fmax(ys, fabs(x[i]));
    for (int i = 0; i < 8; ++i) tn = fmax(tn, fabs(t[i]));
    double eps = 1e-4;
    if (tn * eps > 1e-6 * ys) eps = 1e-6 * ys / tn;
    return eps;
}

// Footprint of a 2x2 Jacobian m (rows: the two outcome coordinates in an orthonormal frame;
// columns: d/du, d/dv) for a circle of radius `size` pixels: semi-axes and the major axis's
// angle from the frame's first axis.
SIRIUS_HD __forceinline__ void ellipse_of(const double m[2][2], double size, float* beam3) {
    const double g00 = size * size * (m[0][0] * m[0][0] + m[0][1] * m[0][1]);
    const double g11 = size * size * (m[1][0] * m[1][0] + m[1][1] * m[1][1]);
    const double g01 = size * size * (m[0][0] * m[1][0] + m[0][1] * m[1][1]);
    const double tr = g00 + g11, det = g00 * g11 - g01 * g01;
    const double disc = sqrt(fmax(0.0, tr * tr - 4.0 * det));
    beam3[0] = (float)sqrt(fmax(0.5 * (tr + disc), 0.0));
    beam3[1] = (float)sqrt(fmax(0.5 * (tr - disc), 0.0));
    beam3[2] = (float)(0.5 * atan2(2.0 * g01, g00 - g11));
}

}  // namespace

template <int FAM, int EMIT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_trace_bundle(const CamConst* __restrict__ cams, int n_cam, TraceParams tp,
                                                     BundleParams bp, BundleOut bo) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= bo.n) return;
    const CamConst& c = cams[find_cam(cams, n_cam, id)];
    const int64_t pix = id - c.offset;
    const double u0 = (double)(pix % c.width) + 0.5, v0 = (double)(pix / c.width) + 0.5;
    // one tangent at a time: a Dual<double, 1> pass per image axis keeps the step's seven stage
    // derivatives in registers (two tangents at once spill); the value parts of both passes are
    // the same arithmetic, so the central ray is the same in each
    using D = Dual<double, 1>;
    const Spacetime<double> s = make_spacetime<double>(c);
    const Spacetime<D> sD = make_spacetime<D>(c);
    double x[4], p[4], t[2][8];
    Der<D> dA[2];  // FSAL derivative of each pass at the step's start
    for (int a = 0; a < 2; ++a) {
        D X[4], P[4];
        init_state_uv<D>(c, sD, a == 0 ? D::seed(u0, 0) : D(u0), a == 1 ? D::seed(v0, 0) : D(v0), X, P);
        eval<D, FAM>(sD, X, P, dA[a]);
        for (int i = 0; i < 4; ++i) {
            x[i] = X[i].v;
            p[i] = P[i].v;
            t[a][i] = X[i].d[0];
            t[a][4 + i] = P[i].d[0];
        }
    }
    // the pass of axis a from the current state: state with that tangent, one DP45 step of h
    auto pass = [&](int a, double hh, D X1[4], D P1[4], Der<D>& d1, D err[4], bool want_err) {
        D X[4], P[4];
        for (int i = 0; i < 4; ++i) {
            X[i] = D(x[i]);
            X[i].d[0] = t[a][i];
            P[i] = D(p[i]);
            P[i].d[0] = t[a][4 + i];
        }
        dp45_step<D, FAM, true>(sD, X, P, dA[a], D(hh), X1, P1, d1, err, want_err);
    };
    double h = tp.h_init;
    if (tp.far_slope > 0) h = fmax(h, fmax(tp.h_max, event_radius(s, x) * tp.far_slope));
    int cls = SIRIUS_RAY_MAX_STEPS;
    double jac[4] = {0, 0, 0, 0};
    uint32_t attempts = 0;
    while (attempts < (uint32_t)tp.max_steps) {
        double hmax = tp.h_max;
        if (tp.far_slope > 0) hmax = fmax(hmax, event_radius(s, x) * tp.far_slope);
        D X1[4], P1[4], err[4];
        Der<D> d1;
        pass(0, h, X1, P1, d1, err, true);
        ++attempts;
        double e = 0;
        for (int i = 0; i < 4; ++i) {
            const double q = err[i].v / (tp.atol + tp.rtol * fabs(X1[i].v));
            e += q * q;
        }
        e = sqrt(0.25 * e);
        if (!(e <= 1.0)) {
            h = next_step(h, e, tp.h_min, hmax);
            if (!sfinite(e) || h <= tp.h_min) { cls = SIRIUS_RAY_FAILED; break; }
            continue;
        }
        double x1[4], p1[4], k0[4], k1[4];
        bool bad = false;
        for (int i = 0; i < 4; ++i) {
            x1[i] = X1[i].v;
            p1[i] = P1[i].v;
            k0[i] = dA[0].k[i].v;
            k1[i] = d1.k[i].v;
            bad |= !sfinite(x1[i]) || !sfinite(p1[i]);
        }
        if (bad) { cls = SIRIUS_RAY_FAILED; break; }
        // thin disk: first crossing inside [r_in, r_out], landed with a DP45 step to the plane
        if constexpr (FAM == 0 && EMIT == SIRIUS_EMISSION_THIN_DISK) {
            if (x[3] * x1[3] < 0) {
                const double th = locate_event(s, EV_Z, 0.0, x, k0, x1, k1, h);
                double xc[4];
                hermite(x, k0, x1, k1, h, th, xc);
                const double rc = kerr_r(s, xc[1], xc[2], xc[3]);
                if (rc >= c.r_in && rc <= c.r_out) {
                    // deviation of the hit point: along the ray onto the plane z = 0
                    for (int a = 0; a < 2; ++a) {
                        D XL[4], PL[4], el[4];
                        Der<D> dl;
                        pass(a, th * h, XL, PL, dl, el, false);
                        const double kz = dl.k[3].v;
                        const double dz = fabs(kz) > 1e-300 ? XL[3].d[0] / kz : 0.0;
                        jac[a] = XL[1].d[0] - dz * dl.k[1].v;
                        jac[2 + a] = XL[2].d[0] - dz * dl.k[2].v;
                    }
                    cls = SIRIUS_RAY_DISK;
                    break;
                }
            }
        }
        const double r1 = event_radius(s, x1);
        if (FAM == 0 && r1 < c.r_cap) { cls = SIRIUS_RAY_CAPTURED; break; }
        // accepted: the second axis's pass, then both tangents and FSAL derivatives advance
        {
            D X2[4], P2[4], e2[4];
            Der<D> d2;
            pass(1, h, X2, P2, d2, e2, false);
            for (int i = 0; i < 4; ++i) {
                t[0][i] = X1[i].d[0];
                t[0][4 + i] = P1[i].d[0];
                t[1][i] = X2[i].d[0];
                t[1][4 + i] = P2[i].d[0];
            }
            dA[0] = d1;
            dA[1] = d2;
        }
        for (int i = 0; i < 4; ++i) { x[i] = x1[i]; p[i] = p1[i]; }
        if (r1 > tp.r_escape) {
            const double radial = (FAM == 1) ? x1[1] * k1[1] : x1[1] * k1[1] + x1[2] * k1[2] + x1[3] * k1[3];
            if (radial > 0) {
                // the escape direction (a closed-form function of the final state and, for the
                // wormhole, of the image point) differentiated along the two deviations
                for (int a = 0; a < 2; ++a) {
                    const double eps = fd_eps(x, p, t[a]);
                    double ang[2][2];
                    for (int sg = 0; sg < 2; ++sg) {
                        const double e3 = sg ? -eps : eps;
                        double xs[4], ps[4];
                        for (int i = 0; i < 4; ++i) {
                            xs[i] = x[i] + e3 * t[a][i];
                            ps[i] = p[i] + e3 * t[a][4 + i];
                        }
                        escape_angles<FAM>(c, s, u0 + (a == 0 ? e3 : 0.0), v0 + (a == 1 ? e3 : 0.0), xs, ps,
                                           ang[sg][0], ang[sg][1]);
                    }
                    double dph = ang[0][1] - ang[1][1];
                    dph = remainder(dph, 2.0 * M_PI);
                    jac[a] = (ang[0][0] - ang[1][0]) / (2.0 * eps);
                    jac[2 + a] = dph / (2.0 * eps);
                }
                cls = SIRIUS_RAY_ESCAPED;
                break;
            }
        }
        h = next_step(h, e, tp.h_min, hmax);
    }
    bo.cls[id] = (uint8_t)cls;
    uint8_t flags = 0;
    float beam[3] = {0.f, 0.f, 0.f};
    if (cls == SIRIUS_RAY_ESCAPED || cls == SIRIUS_RAY_DISK) {
        double m[2][2] = {{jac[0], jac[1]}, {jac[2], jac[3]}};
        if (cls == SIRIUS_RAY_ESCAPED) {
            // (theta, phi) -> the orthonormal sky frame (e_theta, e_phi)
            double th, ph;
            escape_angles<FAM>(c, s, u0, v0, x, p, th, ph);
            m[1][0] *= sin(th);
            m[1][1] *= sin(th);
        }
        ellipse_of(m, bp.size, beam);
        bool fin = true;
        for (int i = 0; i < 4; ++i) fin &= sfinite(jac[i]);
        if (!fin || !(beam[0] >= bp.min_size) || !(beam[0] <= bp.max_ratio * fmax(beam[1], 1e-30f)))
            flags |= SIRIUS_BUNDLE_TERMINATED;
    } else {
        flags |= SIRIUS_BUNDLE_TERMINATED;
    }
    bo.flags[id] = flags;
    for (int i = 0; i < 4; ++i) bo.jac[(int64_t)i * bo.n + id] = (float)jac[i];
    for (int i = 0; i < 3; ++i) bo.beam[(int64_t)i * bo.n + id] = beam[i];
}

hipError_t launch_trace_bundle(int family, int emission, const CamConst* cams, int n_cam, const TraceParams& tp,
                               const BundleParams& bp, const BundleOut& bo, hipStream_t st) {
    if (bo.n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((bo.n + 63) / 64)), block(64);
    const bool disk = emission == SIRIUS_EMISSION_THIN_DISK;
    if (family == 1)
        hipLaunchKernelGGL((k_trace_bundle<1, 0>), grid, block, 0, st, cams, n_cam, tp, bp, bo);
    else if (family == 2)
        hipLaunchKernelGGL((k_trace_bundle<2, 0>), grid, block, 0, st, cams, n_cam, tp, bp, bo);
    else if (disk)
        hipLaunchKernelGGL((k_trace_bundle<0, SIRIUS_EMISSION_THIN_DISK>), grid, block, 0, st, cams, n_cam, tp, bp, bo);
    else
        hipLaunchKernelGGL((k_trace_bundle<0, 0>), grid, block, 0, st, cams, n_cam, tp, bp, bo);
    return hipGetLastError();
}

}  // namespace sirius
