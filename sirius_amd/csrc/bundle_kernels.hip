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
// start as the derivative of the camera's initial state and are carried through every
// accepted DP45 step as the derivative of that step's map (the step with its length held),
// taken by central differences of the step at the perturbed states y +- eps t; the step
// length is chosen on the central ray alone. At the end the derivative of the outcome follows:
// of the escape direction (theta, phi) for an escaped ray (rad per pixel), or of the
// equatorial hit point (X, Y) for a ray ending on the thin disk (M per pixel), the deviation
// there projected along the ray onto the plane.
//
// One thread per ray, FP64. An accepted step costs the central step and four perturbed steps.
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

// Difference scale for a deviation t at state y: a perturbation of 1e-4 pixel, or less where
// the deviation has grown, so that eps |t| stays below 1e-6 of the state's scale.
SIRIUS_HD __forceinline__ double fd_eps(const double x[4], const double p[4], const double t[8]) {
    double ys = 1.0, tn = 0.0;
    for (int i = 1; i < 4; ++i) ys = fmax(ys, fabs(x[i]));
    for (int i = 0; i < 8; ++i) tn = fmax(tn, fabs(t[i]));
    double eps = 1e-4;
    if (tn * eps > 1e-6 * ys) eps = 1e-6 * ys / tn;
    return eps;
}

// The deviations carried through one DP45 step of length h from (x, p): t <- dPhi_h . t.
template <int FAM>
SIRIUS_HD __forceinline__ void step_deviation(const Spacetime<double>& s, const double x[4], const double p[4],
                                              double h, double t[2][8]) {
    for (int a = 0; a < 2; ++a) {
        const double eps = fd_eps(x, p, t[a]);
        double yp[2][8];
        for (int sg = 0; sg < 2; ++sg) {
            const double e = sg ? -eps : eps;
            double xs[4], ps[4], x1[4], p1[4], err[4];
            for (int i = 0; i < 4; ++i) {
                xs[i] = x[i] + e * t[a][i];
                ps[i] = p[i] + e * t[a][4 + i];
            }
            Der<double> d0, d1;
            eval<double, FAM>(s, xs, ps, d0);
            dp45_step<double, FAM, true>(s, xs, ps, d0, h, x1, p1, d1, err, false);
            for (int i = 0; i < 4; ++i) {
                yp[sg][i] = x1[i];
                yp[sg][4 + i] = p1[i];
            }
        }
        for (int i = 0; i < 8; ++i) t[a][i] = (yp[0][i] - yp[1][i]) / (2.0 * eps);
    }
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
__global__ __launch_bounds__(64) void k_trace_bundle(const CamConst* __restrict__ cams, int n_cam, TraceParams tp,
                                                     BundleParams bp, BundleOut bo) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= bo.n) return;
    const CamConst& c = cams[find_cam(cams, n_cam, id)];
    const int64_t pix = id - c.offset;
    const double u0 = (double)(pix % c.width) + 0.5, v0 = (double)(pix / c.width) + 0.5;
    const Spacetime<double> s = make_spacetime<double>(c);
    double x[4], p[4], t[2][8];
    init_state_uv<double>(c, s, u0, v0, x, p);
    {
        // the camera's initial state differentiated in u and v (central, 1e-3 pixel)
        const double e0 = 1e-3;
        for (int a = 0; a < 2; ++a) {
            double xp[4], pp[4], xm[4], pm[4];
            init_state_uv<double>(c, s, u0 + (a == 0 ? e0 : 0.0), v0 + (a == 1 ? e0 : 0.0), xp, pp);
            init_state_uv<double>(c, s, u0 - (a == 0 ? e0 : 0.0), v0 - (a == 1 ? e0 : 0.0), xm, pm);
            for (int i = 0; i < 4; ++i) {
                t[a][i] = (xp[i] - xm[i]) / (2.0 * e0);
                t[a][4 + i] = (pp[i] - pm[i]) / (2.0 * e0);
            }
        }
    }
    Der<double> d;
    eval<double, FAM>(s, x, p, d);
    double h = tp.h_init;
    if (tp.far_slope > 0) h = fmax(h, fmax(tp.h_max, event_radius(s, x) * tp.far_slope));
    int cls = SIRIUS_RAY_MAX_STEPS;
    double jac[4] = {0, 0, 0, 0};
    uint32_t attempts = 0;
    while (attempts < (uint32_t)tp.max_steps) {
        double hmax = tp.h_max;
        if (tp.far_slope > 0) hmax = fmax(hmax, event_radius(s, x) * tp.far_slope);
        double x1[4], p1[4], err[4];
        Der<double> d1;
        dp45_step<double, FAM, true>(s, x, p, d, h, x1, p1, d1, err, true);
        ++attempts;
        double e = 0;
        for (int i = 0; i < 4; ++i) {
            const double q = err[i] / (tp.atol + tp.rtol * fabs(x1[i]));
            e += q * q;
        }
        e = sqrt(0.25 * e);
        if (!(e <= 1.0)) {
            h = next_step(h, e, tp.h_min, hmax);
            if (!sfinite(e) || h <= tp.h_min) { cls = SIRIUS_RAY_FAILED; break; }
            continue;
        }
        bool bad = false;
        for (int i = 0; i < 4; ++i) bad |= !sfinite(x1[i]) || !sfinite(p1[i]);
        if (bad) { cls = SIRIUS_RAY_FAILED; break; }
        // thin disk: first crossing inside [r_in, r_out], landed with a DP45 step to the plane
        if constexpr (FAM == 0 && EMIT == SIRIUS_EMISSION_THIN_DISK) {
            if (x[3] * x1[3] < 0) {
                const double th = locate_event(s, EV_Z, 0.0, x, d.k, x1, d1.k, h);
                double xc[4];
                hermite(x, d.k, x1, d1.k, h, th, xc);
                const double rc = kerr_r(s, xc[1], xc[2], xc[3]);
                if (rc >= c.r_in && rc <= c.r_out) {
                    const double hl = th * h;
                    double xl[4], pl[4], el[4];
                    Der<double> dl;
                    dp45_step<double, FAM, true>(s, x, p, d, hl, xl, pl, dl, el, false);
                    step_deviation<FAM>(s, x, p, hl, t);
                    // deviation of the hit point: along the ray onto the plane z = 0
                    const double kz = dl.k[3];
                    double m[2][2];
                    for (int a = 0; a < 2; ++a) {
                        const double dz = fabs(kz) > 1e-300 ? t[a][3] / kz : 0.0;
                        m[0][a] = t[a][1] - dz * dl.k[1];
                        m[1][a] = t[a][2] - dz * dl.k[2];
                    }
                    jac[0] = m[0][0]; jac[1] = m[0][1]; jac[2] = m[1][0]; jac[3] = m[1][1];
                    cls = SIRIUS_RAY_DISK;
                    break;
                }
            }
        }
        const double r1 = event_radius(s, x1);
        if (FAM == 0 && r1 < c.r_cap) { cls = SIRIUS_RAY_CAPTURED; break; }
        step_deviation<FAM>(s, x, p, h, t);
        for (int i = 0; i < 4; ++i) { x[i] = x1[i]; p[i] = p1[i]; }
        d = d1;
        if (r1 > tp.r_escape) {
            const double radial = (FAM == 1) ? x1[1] * d1.k[1] : x1[1] * d1.k[1] + x1[2] * d1.k[2] + x1[3] * d1.k[3];
            if (radial > 0) {
                // the escape direction differentiated along the two deviations
                for (int a = 0; a < 2; ++a) {
                    const double eps = fd_eps(x, p, t[a]);
                    double ang[2][2];
                    for (int sg = 0; sg < 2; ++sg) {
                        const double e2 = sg ? -eps : eps;
                        double xs[4], ps[4];
                        for (int i = 0; i < 4; ++i) {
                            xs[i] = x[i] + e2 * t[a][i];
                            ps[i] = p[i] + e2 * t[a][4 + i];
                        }
                        escape_angles<FAM>(c, s, u0 + (a == 0 ? e2 : 0.0), v0 + (a == 1 ? e2 : 0.0), xs, ps,
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
