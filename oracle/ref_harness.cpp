// ref_harness.cpp — drives the reference's own CPU code (compiled in place from the reference
// tree by oracle/Makefile, target `ref`) to produce golden vectors and a CPU baseline.
// TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py and bench.py's CPU-baseline
// leg; never linked into the product.
//
// What is the reference's own code here (called, not restated):
//   KerrSchildFamily::evaluate        Sirius.Core/Metric/PHMT100A.h:248-396
//   MorrisThorneFamily::evaluate      Sirius.Core/Metric/PHMT101A.h:285-345
//   WarpDriveFamily::evaluate         Sirius.Core/Metric/PHMT102A.h:515-578
//   PinholeCamera::generateRay        Sirius.Core/Camera/CMBS001A.h:120-185
//   Coordinates::BLToCartesian        Sirius.Core/Coordinate/PHCT002A.h:129-140
//   TensorOps::normalizeNull          Sirius.Core/Tensor/MTTN001A.cpp:171-237
//   Geodesic::integrateStepRK45       Sirius.Core/Geodesic/PHGD001A.cpp:530-641
//   AccretionDiskD (ISCO, Page-Thorne quadrature)  Sirius.Core/Disk/PHAD001A.h:88-314
//   SymplecticIntegratorD             Sirius.Core/Symplectic/PHSI001A.h:136-455
//   StarfieldGenerator                Sirius.Core/Environment/PHSF001A.h:126-272
//   Spectral::blackbodyToRGB, applyLimbDarkening  Sirius.Core/Spectral/PHSP001A.h:181-205, 306-322
//   ColorModes::applyColorMode (TrueColor)        Sirius.Core/Spectral/PHSC001A.h:343-366
// What the harness supplies itself: GeodesicTracer::initializeLightray is private, so its
// steps (GTRC001A.cpp:15-134, including the float casts) are repeated here; the step loop
// retries rejected steps (GeodesicTracer::trace ends a ray on the first rejection,
// GTRC001A.cpp:269-277, which at rtol 1e-8 ends almost every ray as MAX_STEPS) and applies
// the benchmark's termination radii on the Kerr r.
#include "GTRC001A.h"
#include "PHMT101A.h"
#include "PHMT102A.h"
#include "PHAD001A.h"
#include "PHSI001A.h"
#include "Environment/PHSF001A.h"
#include "PHSC001A.h"
#include "PHSP001A.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace Sirius;

namespace {

void copy_out(const Metric4D& g, const Tensor<Dual<double>, 4, 4, 4>& dg, double* G, double* DG) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) G[i * 4 + j] = g(i, j).real;
    for (int l = 0; l < 4; ++l)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) DG[l * 16 + i * 4 + j] = dg(l, i, j).real;
}

CameraConfig cam_cfg(double r, double theta, double phi, double fov, int w, int h) {
    CameraConfig c;
    c.r = r;
    c.theta = theta;
    c.phi = phi;
    c.fov = static_cast<float>(fov);
    c.width = w;
    c.height = h;
    return c;
}

// GTRC001A.cpp:15-106, step by step (the method is private to GeodesicTracer).
Lightray init_ray(IMetric* metric, const CameraRay& cr, float h0) {
    Lightray ray{};
    double t = cr.origin(0), r = cr.origin(1), th = cr.origin(2), ph = cr.origin(3);
    th = std::clamp(th, 0.01, M_PI - 0.01);
    Coordinates::Vec4BL bl(t, r, th, ph);
    Coordinates::Vec4Cart c = Coordinates::BLToCartesian(bl);
    ray.position(0) = static_cast<float>(c.t);
    ray.position(1) = static_cast<float>(c.x);
    ray.position(2) = static_cast<float>(c.y);
    ray.position(3) = static_cast<float>(c.z);
    double st = std::sin(th), ct = std::cos(th), sp = std::sin(ph), cp = std::cos(ph);
    double vr = cr.direction(1), vth = cr.direction(2) / r, vph = cr.direction(3) / (r * st);
    ray.velocity(1) = static_cast<float>(vr * st * cp + r * vth * ct * cp - r * vph * st * sp);
    ray.velocity(2) = static_cast<float>(vr * st * sp + r * vth * ct * sp + r * vph * st * cp);
    ray.velocity(3) = static_cast<float>(vr * ct - r * vth * st);
    Metric4D g;
    Tensor<Dual<double>, 4, 4, 4> dg;
    metric->evaluate(ray.position, g, dg);
    Vec4 v;
    v(0) = 0.0;
    for (int i = 1; i < 4; ++i) v(i) = ray.velocity(i);
    Vec4 n = TensorOps::normalizeNull(v, g);
    for (int i = 0; i < 4; ++i) ray.velocity(i) = static_cast<float>(n(i));
    ray.step_size = h0;
    return ray;
}

}  // namespace

extern "C" {

void ref_ks_evaluate(double M, double a, double Q, double L, const double* x, double* G, double* DG) {
    KerrSchildFamily m(KerrSchildParams{M, a, Q, L});
    Vec4 p;
    for (int i = 0; i < 4; ++i) p(i) = x[i];
    Metric4D g;
    Tensor<Dual<double>, 4, 4, 4> dg;
    m.evaluate(p, g, dg);
    copy_out(g, dg, G, DG);
}

void ref_mt_evaluate(double b0, const double* x, double* G, double* DG) {
    MorrisThorneFamily m(MorrisThorneParams::Ellis(b0));
    Vec4 p;
    for (int i = 0; i < 4; ++i) p(i) = x[i];
    Metric4D g;
    Tensor<Dual<double>, 4, 4, 4> dg;
    m.evaluate(p, g, dg);
    copy_out(g, dg, G, DG);
}

void ref_warp_evaluate(double v, double R, double sigma, const double* x, double* G, double* DG) {
    WarpDriveFamily m(WarpDriveParams::Alcubierre(v, R, sigma));
    Vec4 p;
    for (int i = 0; i < 4; ++i) p(i) = x[i];
    Metric4D g;
    Tensor<Dual<double>, 4, 4, 4> dg;
    m.evaluate(p, g, dg);
    copy_out(g, dg, G, DG);
}

double ref_ks_radius(double a, double x, double y, double z) {
    KerrSchildFamily m(KerrSchildParams{1.0, a, 0.0, 0.0});
    return m.computeKerrRadius(x, y, z);
}

// PinholeCamera::generateRay: origin (t, r, th, ph) and direction (0, v_r, r v_th, r sin v_ph).
void ref_camera_ray(double r, double theta, double phi, double fov, int w, int h, int px, int py,
                    double* origin, double* dir) {
    PinholeCamera cam(cam_cfg(r, theta, phi, fov, w, h));
    CameraRay cr = cam.generateRay(px, py);
    for (int i = 0; i < 4; ++i) {
        origin[i] = cr.origin(i);
        dir[i] = cr.direction(i);
    }
}

// Initial Cartesian state (x, k) of GTRC001A.cpp:15-106 for one pixel.
void ref_init_state(double M, double a, double Q, double L, double r, double theta, double phi,
                    double fov, int w, int h, int px, int py, double* x, double* k) {
    KerrSchildFamily m(KerrSchildParams{M, a, Q, L});
    PinholeCamera cam(cam_cfg(r, theta, phi, fov, w, h));
    Lightray ray = init_ray(&m, cam.generateRay(px, py), 0.01f);
    for (int i = 0; i < 4; ++i) {
        x[i] = ray.position(i);
        k[i] = ray.velocity(i);
    }
}

double ref_normalize_null(const double* G, const double* k) {
    Metric4D g;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) g(i, j) = Dual<double>(G[i * 4 + j]);
    Vec4 v;
    for (int i = 0; i < 4; ++i) v(i) = k[i];
    return TensorOps::normalizeNull(v, g)(0);
}

// One call of Geodesic::integrateStepRK45. Returns 1 accepted, 0 rejected, -1 terminated.
int ref_rk45_step(double M, double a, double Q, double L, double* x, double* k, double* h,
                  double rtol, double atol, double hmin, double hmax) {
    KerrSchildFamily m(KerrSchildParams{M, a, Q, L});
    IntegratorConfig cfg;
    cfg.rel_tolerance = static_cast<float>(rtol);
    cfg.abs_tolerance = static_cast<float>(atol);
    cfg.min_step = static_cast<float>(hmin);
    cfg.max_step = static_cast<float>(hmax);
    Lightray ray{};
    for (int i = 0; i < 4; ++i) {
        ray.position(i) = x[i];
        ray.velocity(i) = k[i];
    }
    ray.step_size = static_cast<float>(*h);
    bool ok = Geodesic::integrateStepRK45(ray, &m, cfg);
    for (int i = 0; i < 4; ++i) {
        x[i] = ray.position(i);
        k[i] = ray.velocity(i);
    }
    *h = ray.step_size;
    if (ray.terminated) return -1;
    return ok ? 1 : 0;
}

// Trace pixels with the reference's integrator. Per ray out: cls (0 esc, 1 capt, 2 disk,
// 3 max, 5 failed), steps (attempts), final x[4], k[4], disk r (BL r of the linear crossing
// point, GTRC001A.h:404-431 semantics), affine length.
void ref_trace(double M, double a, double Q, double L, double r_obs, double theta, double phi,
               double fov, int w, int h, const int64_t* pix, int64_t n, double rtol, double atol,
               double h0, double hmin, double hmax, int max_steps, double r_cap, double r_esc,
               double disk_in, double disk_out, int use_disk, int threads, int32_t* cls,
               int32_t* steps, double* xf, double* kf, double* disk_r, double* lam) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
        KerrSchildFamily m(KerrSchildParams{M, a, Q, L});
        PinholeCamera cam(cam_cfg(r_obs, theta, phi, fov, w, h));
        IntegratorConfig cfg;
        cfg.rel_tolerance = static_cast<float>(rtol);
        cfg.abs_tolerance = static_cast<float>(atol);
        cfg.min_step = static_cast<float>(hmin);
        cfg.max_step = static_cast<float>(hmax);
        cfg.initial_step = static_cast<float>(h0);
        int px = static_cast<int>(pix[i] % w), py = static_cast<int>(pix[i] / w);
        Lightray ray = init_ray(&m, cam.generateRay(px, py), cfg.initial_step);
        int c = 3, s = 0;
        double dr = 0, lsum = 0;
        while (s < max_steps) {
            Vec4 prev = ray.position;
            float hh = ray.step_size;
            bool ok = Geodesic::integrateStepRK45(ray, &m, cfg);
            ++s;
            if (ray.terminated) { c = 5; break; }
            if (!ok) continue;  // rejected: retry with the reduced step
            lsum += hh;
            double X = ray.position(1), Y = ray.position(2), Z = ray.position(3);
            double rk = m.computeKerrRadius(X, Y, Z);
            if (use_disk && prev(3) * Z < 0) {
                double al = -prev(3) / (Z - prev(3));
                double xc = prev(1) + al * (X - prev(1)), yc = prev(2) + al * (Y - prev(2));
                double rc = std::sqrt(std::max(xc * xc + yc * yc - a * a, 0.0));
                if (rc >= disk_in && rc <= disk_out) { c = 2; dr = rc; break; }
            }
            if (rk < r_cap) { c = 1; break; }
            if (rk > r_esc) {
                double vr = X * ray.velocity(1) + Y * ray.velocity(2) + Z * ray.velocity(3);
                if (vr > 0) { c = 0; break; }
            }
        }
        cls[i] = c;
        steps[i] = s;
        disk_r[i] = dr;
        lam[i] = lsum;
        for (int j = 0; j < 4; ++j) {
            xf[4 * i + j] = ray.position(j);
            kf[4 * i + j] = ray.velocity(j);
        }
    }
}

double ref_isco(double a) { return AccretionDiskD::computeISCO(a); }

// Page-Thorne Q(r) = -Omega_r / (E - Omega L)^2 * Int (E - Omega L) L_r dr as the reference's
// 16-point quadrature computes it: fullPageThorneFlux(r) divided by its Newtonian prefactor.
double ref_pt_q(double a, double r) {
    AccretionDiskD::Config cfg;
    cfg.M = 1.0;
    cfg.a_star = a;
    cfg.r_outer = 1e4;
    AccretionDiskD d(cfg);
    double F = d.fullPageThorneFlux(r);
    double Mkg = cfg.M * Constants::Physical::M_SUN;
    double GM = Constants::Physical::G_NEWTON * Mkg;
    double rs = 2 * GM / (Constants::Physical::c_LIGHT * Constants::Physical::c_LIGHT);
    double mdot = cfg.Mdot * Constants::Physical::M_SUN / (365.25 * 24 * 3600);
    double rp = r * rs / 2;
    double Fn = 3 * GM * mdot / (8 * M_PI * rp * rp * rp);
    return F / Fn;
}

// E, L, Omega of circular equatorial orbits (PHAD001A.h:112-142).
void ref_circular(double a, double r, double* E, double* Lz, double* Om) {
    *E = AccretionDiskD::specificEnergy(r, 1.0, a);
    *Lz = AccretionDiskD::specificAngularMomentum(r, 1.0, a);
    *Om = AccretionDiskD::angularVelocity(r, 1.0, a);
}

// Reference symplectic integrator (BL Kerr) conservation run: returns max |H| and the final
// E and L_z drift after n steps of size h from (r, theta, p_r, p_theta, E, L).
void ref_symplectic_run(double a, double r, double th, double pr, double pth, double E, double Lz,
                        int n, double h, double* maxH, double* dE, double* dL) {
    KerrMetricD km(1.0, a, 0.0);
    SymplecticIntegratorD::Config cfg;
    SymplecticIntegratorD integ(&km, cfg);
    GeodesicStateD s;
    s.x = Vec4d(0.0, r, th, 0.0);
    s.k = Vec4d(-E, pr, pth, Lz);
    auto st = integ.testConservation(s, n, h);
    *maxH = st.maxHError;
    *dE = st.finalE - st.initialE;
    *dL = st.finalLz - st.initialLz;
}

// StarfieldGenerator::generate (PHSF001A.h:133-151): the catalogue, 8 floats per star in
// StarEntry order; returns the number of stars (at most cap are written).
int ref_starfield_generate(uint32_t count, float dmin, float dmax, float maglim, uint32_t seed, float* out,
                           int cap) {
    StarfieldConfig c;
    c.star_count = count;
    c.min_distance_pc = dmin;
    c.max_distance_pc = dmax;
    c.magnitude_limit = maglim;
    c.seed = seed;
    StarfieldGenerator g(c);
    std::vector<StarEntry> stars = g.generate();
    for (size_t i = 0; i < stars.size() && (int)i < cap; ++i) std::memcpy(out + 8 * i, &stars[i], 8 * sizeof(float));
    return (int)stars.size();
}

// StarfieldGenerator::sampleStarfield (PHSF001A.h:153-201) at m unit directions (camera at the
// origin, no parallax offset): rgb, 3 floats per direction.
void ref_starfield_sample(float aperture_mm, float focus_pc, float brightness, int dof, const float* stars, int n,
                          const float* dirs, int m, float* rgb) {
    StarfieldConfig c;
    c.aperture_mm = aperture_mm;
    c.focus_distance_pc = focus_pc;
    c.brightness_scale = brightness;
    c.enable_dof = dof != 0;
    StarfieldGenerator g(c);
    std::vector<StarEntry> cat(n);
    for (int i = 0; i < n; ++i) std::memcpy(&cat[i], stars + 8 * i, 8 * sizeof(float));
    for (int j = 0; j < m; ++j)
        g.sampleStarfield(dirs[3 * j], dirs[3 * j + 1], dirs[3 * j + 2], 0.f, 0.f, 0.f, cat, rgb[3 * j], rgb[3 * j + 1],
                          rgb[3 * j + 2]);
}

// Spectral::blackbodyToRGB (PHSP001A.h:181-205) at n temperatures (K): rgb, 3 floats each.
void ref_blackbody_rgb(const double* T, int n, float* rgb) {
    for (int i = 0; i < n; ++i) {
        Spectral::RGB c = Spectral::blackbodyToRGB(T[i]);
        rgb[3 * i] = c.r;
        rgb[3 * i + 1] = c.g;
        rgb[3 * i + 2] = c.b;
    }
}

// ColorModes::applyColorMode(TrueColor, T_emit, g, intensity) (PHSC001A.h:343-366), n cases.
void ref_true_color(const float* T_emit, const float* g, const float* intensity, int n, float* rgb) {
    for (int i = 0; i < n; ++i) {
        Spectral::RGB c = ColorModes::applyColorMode(ColorModes::Mode::TrueColor, T_emit[i], g[i], intensity[i], nullptr);
        rgb[3 * i] = c.r;
        rgb[3 * i + 1] = c.g;
        rgb[3 * i + 2] = c.b;
    }
}

// Spectral::applyLimbDarkening (PHSP001A.h:306-322) of colours rgb_in at cosines cos_i, n cases.
void ref_limb_darkening(const float* rgb_in, const double* cos_i, int n, float* rgb) {
    for (int i = 0; i < n; ++i) {
        Spectral::RGB c = Spectral::applyLimbDarkening(Spectral::RGB(rgb_in[3 * i], rgb_in[3 * i + 1], rgb_in[3 * i + 2]),
                                                       cos_i[i]);
        rgb[3 * i] = c.r;
        rgb[3 * i + 1] = c.g;
        rgb[3 * i + 2] = c.b;
    }
}

}  // extern "C"
