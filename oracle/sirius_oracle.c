/*
 * sirius_oracle.c — CPU FP64 oracle of the geodesic-tracing path. TEST INFRASTRUCTURE ONLY:
 * the checker for the HIP backend, loaded by tests/, the smoke entry and bench.py's CPU
 * baseline leg, never by the product library.
 *
 * Reference lines followed (paths relative to the reference's src/):
 *   Kerr-Schild family metric        Sirius.Core/Metric/PHMT100A.h:198-396
 *   Morris-Thorne (Ellis) metric      Sirius.Core/Metric/PHMT101A.h:195-345 (restated in the
 *                                     two-sided proper-length coordinate l, r^2 = l^2 + b0^2)
 *   Alcubierre metric                Sirius.Core/Metric/PHMT102A.h:485-578 (static bubble)
 *   Hamiltonian DP45 step            Sirius.Core/Geodesic/PHGD001A.cpp:16-79,159-202,471-641
 *   null re-normalisation            Sirius.Core/Tensor/MTTN001A.cpp:171-237
 *   pinhole camera                   Sirius.Core/Camera/CMBS001A.h:120-185
 *   camera ray -> Cartesian state    Sirius.Render/Integration/GTRC001A.cpp:15-134,
 *                                    Sirius.Core/Coordinate/PHCT002A.h:129-140
 *   termination / disk crossing      Sirius.Render/Integration/GTRC001A.cpp:264-556,
 *                                    GTRC001A.h:365-431
 *   ISCO, circular-orbit E, L, Omega Sirius.Core/Disk/PHAD001A.h:88-142
 *   Page-Thorne flux                 Sirius.Core/Disk/PHAD001A.h:234-373 (closed form of the
 *                                    same integral, Page & Thorne 1974)
 *   Yoshida weights                  Sirius.Core/Symplectic/PHSI001A.h:43-80
 *
 * Departures from the reference, all deliberate and listed in the design notes: pure FP64
 * (no float casts), termination on the Kerr r (not the Euclidean R), exact event location by
 * a landing step (not linear interpolation), first disk crossing terminates, rejected steps
 * are retried (the reference tracer ends a ray on its first rejected step, GTRC001A.cpp:273).
 */
#include "sirius_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ========================================================================================
 * Forward-mode dual numbers with three partials (d/dx, d/dy, d/dz).
 * Restates the arithmetic of Dual<T> (Sirius.Core/Autodiff/MTDL001A.h:19-368) with a
 * vector of tangents, so one evaluation yields a value and its spatial gradient.
 * ======================================================================================== */
typedef struct { double v, d[3]; } dual3;

static dual3 dc(double v) { dual3 r = {v, {0, 0, 0}}; return r; }
static dual3 dvar(double v, int i) { dual3 r = dc(v); r.d[i] = 1.0; return r; }
static dual3 dadd(dual3 a, dual3 b) { dual3 r; r.v = a.v + b.v; for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
static dual3 dsub(dual3 a, dual3 b) { dual3 r; r.v = a.v - b.v; for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
static dual3 dmul(dual3 a, dual3 b) { dual3 r; r.v = a.v * b.v; for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
static dual3 dscale(dual3 a, double s) { dual3 r; r.v = a.v * s; for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * s; return r; }
static dual3 ddiv(dual3 a, dual3 b) {
    dual3 r; r.v = a.v / b.v;
    for (int i = 0; i < 3; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
static dual3 dsqrt(dual3 a) {
    dual3 r; r.v = sqrt(a.v);
    for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] / (2.0 * r.v);
    return r;
}
static dual3 dmax_c(dual3 a, double floor_v) { return a.v < floor_v ? dc(floor_v) : a; }

/* ========================================================================================
 * Metrics
 * ======================================================================================== */

/* Kerr radius from r^4 - (R^2 - a^2) r^2 - a^2 z^2 = 0 (PHMT100A.h:198-216). */
static dual3 ks_radius_dual(double a, dual3 x, dual3 y, dual3 z) {
    double a2 = a * a;
    dual3 R2 = dadd(dadd(dmul(x, x), dmul(y, y)), dmul(z, z));
    dual3 w = dsub(R2, dc(a2));
    dual3 disc = dadd(dmul(w, w), dscale(dmul(z, z), 4.0 * a2));
    disc = dmax_c(disc, 1e-20);
    dual3 r2 = dscale(dadd(w, dsqrt(disc)), 0.5);
    r2 = dmax_c(r2, 1e-10);
    return dsqrt(r2);
}

double oracle_kerr_radius(const SiriusHipMetric* m, double x, double y, double z) {
    if (m->family != SIRIUS_METRIC_KERR_SCHILD) return sqrt(x * x + y * y + z * z);
    return ks_radius_dual(m->a, dc(x), dc(y), dc(z)).v;
}

double oracle_horizon(const SiriusHipMetric* m) {
    /* r_+ = M + sqrt(M^2 - a^2 - Q^2), PHMT100A.h:402-415 */
    if (m->family != SIRIUS_METRIC_KERR_SCHILD) return 0.0;
    double d = m->M * m->M - m->a * m->a - m->Q * m->Q;
    if (d < 0) return 0.0;
    return m->M + sqrt(d);
}

/* g_{mu nu} = eta + H l_mu l_nu, g_tt -= Lambda R^2 / 3 (PHMT100A.h:248-396). The gradient
 * comes from dual numbers, not from the reference's hand derivatives; the tests pin the
 * two against each other and against finite differences. */
static void ks_metric(const SiriusHipMetric* m, const double X[4], double g[16], double dg[64]) {
    double a = m->a, M = m->M, Q2 = m->Q * m->Q, a2 = a * a;
    dual3 x = dvar(X[1], 0), y = dvar(X[2], 1), z = dvar(X[3], 2);
    dual3 r = ks_radius_dual(a, x, y, z);
    dual3 r2 = dmul(r, r);
    dual3 den = dadd(r2, dc(a2));
    dual3 l[4];
    l[0] = dc(1.0);
    l[1] = ddiv(dadd(dmul(r, x), dscale(y, a)), den);
    l[2] = ddiv(dsub(dmul(r, y), dscale(x, a)), den);
    l[3] = ddiv(z, r);
    dual3 num = dmul(dsub(dscale(r, 2.0 * M), dc(Q2)), r2);
    dual3 fden = dadd(dmul(r2, r2), dscale(dmul(z, z), a2));
    dual3 H = ddiv(num, fden);
    memset(dg, 0, 64 * sizeof(double));
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) {
            dual3 t = dmul(H, dmul(l[mu], l[nu]));
            double eta = (mu == nu) ? (mu == 0 ? -1.0 : 1.0) : 0.0;
            g[mu * 4 + nu] = eta + t.v;
            for (int i = 0; i < 3; ++i) dg[(i + 1) * 16 + mu * 4 + nu] = t.d[i];
        }
    if (m->Lambda != 0.0) {
        double R2 = X[1] * X[1] + X[2] * X[2] + X[3] * X[3];
        g[0] -= m->Lambda * R2 / 3.0;
        for (int i = 0; i < 3; ++i) dg[(i + 1) * 16] -= 2.0 * m->Lambda * X[i + 1] / 3.0;
    }
}

/* Ellis wormhole in the proper-length coordinate, reduced to the orbital plane:
 * coordinates (t, l, psi, w), ds^2 = -dt^2 + dl^2 + (l^2 + b0^2) dpsi^2 + dw^2. Same
 * spacetime as PHMT101A.h:285-345 with b(r) = b0^2/r and Phi = 0, but regular through the
 * throat (r = sqrt(l^2 + b0^2)), so rays can cross to the far side. */
static void mt_metric(const SiriusHipMetric* m, const double X[4], double g[16], double dg[64]) {
    double b0 = m->throat_b0, l = X[1];
    memset(g, 0, 16 * sizeof(double));
    memset(dg, 0, 64 * sizeof(double));
    g[0] = -1.0;
    g[5] = 1.0;
    g[10] = l * l + b0 * b0;
    g[15] = 1.0;
    dg[1 * 16 + 2 * 4 + 2] = 2.0 * l;
}

/* Alcubierre: ds^2 = -(1 - v^2 f^2) dt^2 - 2 v f dx dt + dx^2 + dy^2 + dz^2, bubble at the
 * origin (PHMT102A.h:485-578 with x_s = 0, as the reference's static evaluation). */
static void warp_metric(const SiriusHipMetric* m, const double X[4], double g[16], double dg[64]) {
    double v = m->warp_v, R = m->warp_R, s = m->warp_sigma;
    dual3 x = dvar(X[1], 0), y = dvar(X[2], 1), z = dvar(X[3], 2);
    dual3 rs = dsqrt(dmax_c(dadd(dadd(dmul(x, x), dmul(y, y)), dmul(z, z)), 1e-20));
    double tp = tanh(s * (rs.v + R)), tm = tanh(s * (rs.v - R)), tr = tanh(s * R);
    double f = (tp - tm) / (2.0 * tr);
    double dfdr = s * ((1.0 - tp * tp) - (1.0 - tm * tm)) / (2.0 * tr);
    memset(g, 0, 16 * sizeof(double));
    memset(dg, 0, 64 * sizeof(double));
    g[0] = -(1.0 - v * v * f * f);
    g[1] = g[4] = -v * f;
    g[5] = g[10] = g[15] = 1.0;
    for (int i = 0; i < 3; ++i) {
        double df = dfdr * rs.d[i];
        dg[(i + 1) * 16 + 0] = 2.0 * v * v * f * df;
        dg[(i + 1) * 16 + 1] = dg[(i + 1) * 16 + 4] = -v * df;
    }
}

/* General 4x4 inverse by Gauss-Jordan with partial pivoting (independent of the closed
 * forms the device code uses). */
static void inv4(const double A[16], double Ai[16]) {
    double m[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) m[i][j] = j < 4 ? A[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int i = c + 1; i < 4; ++i) if (fabs(m[i][c]) > fabs(m[p][c])) p = i;
        if (p != c) for (int j = 0; j < 8; ++j) { double t = m[c][j]; m[c][j] = m[p][j]; m[p][j] = t; }
        double d = m[c][c];
        for (int j = 0; j < 8; ++j) m[c][j] /= d;
        for (int i = 0; i < 4; ++i) {
            if (i == c) continue;
            double f = m[i][c];
            for (int j = 0; j < 8; ++j) m[i][j] -= f * m[c][j];
        }
    }
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) Ai[i * 4 + j] = m[i][j + 4];
}

void oracle_metric(const SiriusHipMetric* m, const double x[4], double g[16], double ginv[16],
                   double dg[64]) {
    switch (m->family) {
        case SIRIUS_METRIC_MORRIS_THORNE: mt_metric(m, x, g, dg); break;
        case SIRIUS_METRIC_ALCUBIERRE: warp_metric(m, x, g, dg); break;
        default: ks_metric(m, x, g, dg); break;
    }
    if (ginv) inv4(g, ginv);
}

/* dx^mu = g^{mu nu} p_nu; dp_mu = 1/2 d_mu g_{rs} k^r k^s (PHGD001A.cpp:159-202). */
double oracle_rhs(const SiriusHipMetric* m, const double x[4], const double p[4], double dx[4],
                  double dp[4]) {
    double g[16], gi[16], dg[64];
    oracle_metric(m, x, g, gi, dg);
    for (int mu = 0; mu < 4; ++mu) {
        double s = 0;
        for (int nu = 0; nu < 4; ++nu) s += gi[mu * 4 + nu] * p[nu];
        dx[mu] = s;
    }
    for (int lam = 0; lam < 4; ++lam) {
        double s = 0;
        for (int r = 0; r < 4; ++r)
            for (int q = 0; q < 4; ++q) s += dg[lam * 16 + r * 4 + q] * dx[r] * dx[q];
        dp[lam] = 0.5 * s;
    }
    double pk = 0;
    for (int mu = 0; mu < 4; ++mu) pk += p[mu] * dx[mu];
    return pk;
}

/* ========================================================================================
 * Closed forms: horizon, ISCO, photon orbits, Page-Thorne
 * ======================================================================================== */

/* Bardeen-Press-Teukolsky ISCO (PHAD001A.h:88-105); a < 0 is retrograde. M = 1. */
double oracle_isco(double a_star) {
    double a = fabs(a_star);
    if (a > 1) a = 1;
    double z1 = 1 + cbrt(1 - a * a) * (cbrt(1 + a) + cbrt(1 - a));
    double z2 = sqrt(3 * a * a + z1 * z1);
    double s = sqrt((3 - z1) * (3 + z1 + 2 * z2));
    return a_star >= 0 ? 3 + z2 - s : 3 + z2 + s;
}

/* Circular photon orbit r = 2(1 + cos(2/3 acos(-+a))) (GTRC001A.cpp:179-184), M = 1. */
double oracle_photon_orbit(double a, int prograde) {
    double arg = prograde ? -a : a;
    if (arg > 1) arg = 1;
    if (arg < -1) arg = -1;
    return 2.0 * (1.0 + cos((2.0 / 3.0) * acos(arg)));
}

/* Page & Thorne (1974) zero-torque flux factor in closed form (M = 1):
 * f = -Omega_r (E - Omega L)^-2 Int_{r_ms}^{r} (E - Omega L) L_r dr, the integral that
 * PHAD001A.h:234-314 evaluates by 16-point Gauss-Legendre. x = sqrt(r). */
double oracle_page_thorne_f(double r, double a) {
    double rms = oracle_isco(a);
    if (r <= rms) return 0.0;
    double x = sqrt(r), x0 = sqrt(rms);
    double ac = acos(a);
    double x1 = 2.0 * cos((ac - M_PI) / 3.0);
    double x2 = 2.0 * cos((ac + M_PI) / 3.0);
    double x3 = -2.0 * cos(ac / 3.0);
    double br = x - x0 - 1.5 * a * log(x / x0)
              - 3.0 * (x1 - a) * (x1 - a) / (x1 * (x1 - x2) * (x1 - x3)) * log((x - x1) / (x0 - x1))
              - 3.0 * (x2 - a) * (x2 - a) / (x2 * (x2 - x1) * (x2 - x3)) * log((x - x2) / (x0 - x2))
              - 3.0 * (x3 - a) * (x3 - a) / (x3 * (x3 - x1) * (x3 - x2)) * log((x - x3) / (x0 - x3));
    return 1.5 / (x * x * (x * x * x - 3.0 * x + 2.0 * a)) * br;
}

/* max over [r_in, r_out] of f(r)/r, the flux up to Mdot/(4 pi). Dense scan + golden search. */
double oracle_disk_tmax_f_over_r(double a, double r_in, double r_out) {
    double rms = oracle_isco(a);
    double lo = r_in > rms ? r_in : rms;
    if (r_out <= lo) return 0.0;
    int n = 4096;
    double best = -1, rb = lo;
    for (int i = 0; i <= n; ++i) {
        double r = lo + (r_out - lo) * i / n;
        double v = oracle_page_thorne_f(r, a) / r;
        if (v > best) { best = v; rb = r; }
    }
    double h = (r_out - lo) / n;
    double A = rb - h > lo ? rb - h : lo, B = rb + h < r_out ? rb + h : r_out;
    const double gr = 0.6180339887498949;
    for (int it = 0; it < 200; ++it) {
        double c = B - gr * (B - A), d = A + gr * (B - A);
        if (oracle_page_thorne_f(c, a) / c > oracle_page_thorne_f(d, a) / d) B = d; else A = c;
    }
    double rr = 0.5 * (A + B);
    double v = oracle_page_thorne_f(rr, a) / rr;
    return v > best ? v : best;
}

double oracle_disk_temperature_ratio(double r, double a, double r_in, double r_out) {
    if (r < r_in || r > r_out) return 0.0;
    double fmax = oracle_disk_tmax_f_over_r(a, r_in, r_out);
    double f = oracle_page_thorne_f(r, a) / r;
    if (fmax <= 0 || f <= 0) return 0.0;
    return pow(f / fmax, 0.25);
}

void oracle_bin_centres(int n, double x_min, double x_max, double* c) {
    for (int k = 0; k < n; ++k) c[k] = x_min * pow(x_max / x_min, (k + 0.5) / n);
}

/* ========================================================================================
 * Options
 * ======================================================================================== */
void oracle_default_opts(SiriusHipTraceOpts* o, int precision) {
    memset(o, 0, sizeof(*o));
    o->integrator = SIRIUS_INTEGRATOR_DP45;
    o->precision = precision;
    o->emission = SIRIUS_EMISSION_THIN_DISK;
    o->schedule = precision == SIRIUS_PRECISION_FP64 ? SIRIUS_SCHEDULE_GRID : SIRIUS_SCHEDULE_PERSISTENT;
    /* h0 = 0.01 and h_min = 1e-8 as PHCN001A.h:74,77; rtol 1e-8, atol 1e-10 and h_max 2 are the
       oracle's FP64 settings */
    o->rtol = 1e-8;
    o->atol = 1e-10;
    o->h_init = 0.01;
    o->h_min = 1e-8;
    o->h_max = 2.0;
    o->far_slope = 0.0;
    o->symplectic_h = 0.01;
    o->symplectic_order = 4;
    o->max_steps = 20000;   /* SNRS001A.cpp:106 */
    o->r_escape = 1000.0;
    o->horizon_factor = 1.01;
    o->n_bins = 16;
    o->x_min = 0.1;
    o->x_max = 10.0;
    o->escalate_margin = 0.0;
    o->escalate_band = 0.0;
    o->surrogate_band = 0.5;
    o->surrogate_pmin = 0.99;
    o->audit_fraction = 0.01;
    o->audit_seed = 0x51121007ULL;
    o->vol_j0 = 1.0;
    o->vol_h_over_r = 0.1;
    o->vol_kappa = 0.3;
    o->vol_r_min = 2.0;
    o->vol_r_max = 50.0;
    o->vol_tau_max = 10.0;
    o->escalate_disk_edge = 0.0;
    o->escalate_rtol = 0.0;
    o->critical_band = 0.5;
    o->step_radial_exp = 0.0;
    o->surrogate_err_max = 5e-5;
    o->surrogate_margin = 2.5;
    o->precision_audit = 0.0;
    if (precision == SIRIUS_PRECISION_FP32_ESCALATE) {
        o->rtol = 1e-6;
        o->atol = 1e-8;
        o->far_slope = 0.25;
        o->step_radial_exp = 2.0;
    }
}

/* ========================================================================================
 * Camera (CMBS001A.h:120-185) and initial state (GTRC001A.cpp:15-134), in FP64
 * ======================================================================================== */
/* Offset (in pixels) added to every ray's image point; 0 except while a test differentiates
 * traces with respect to the image coordinates (oracle_set_subpixel). */
static double g_sub_u = 0.0, g_sub_v = 0.0;

void oracle_set_subpixel(double du, double dv) {
    g_sub_u = du;
    g_sub_v = dv;
}

static void camera_direction(const SiriusHipCamera* c, int px, int py, double n_sph[3]) {
    double aspect = (double)c->width / c->height;
    double th = tan(c->fov_deg * M_PI / 360.0);
    double sx = (2.0 * (px + 0.5 + g_sub_u) / c->width - 1.0) * aspect;
    double sy = 1.0 - 2.0 * (py + 0.5 + g_sub_v) / c->height;
    double dx = sx * th, dy = sy * th, dz = -1.0;
    double cy = cos(c->yaw), sy_ = sin(c->yaw), cp = cos(c->pitch), sp = sin(c->pitch);
    double cr = cos(c->roll), sr = sin(c->roll);
    double rx = dx * cr - dy * sr, ry = dx * sr + dy * cr;
    double ry2 = ry * cp - dz * sp, rz = ry * sp + dz * cp;
    double rx2 = rx * cy + rz * sy_, rz2 = -rx * sy_ + rz * cy;
    double len = sqrt(rx2 * rx2 + ry2 * ry2 + rz2 * rz2);
    n_sph[0] = rz2 / len;   /* along e_r      */
    n_sph[1] = -ry2 / len;  /* along e_theta  */
    n_sph[2] = rx2 / len;   /* along e_phi    */
}

/* Future-directed root of g_00 k0^2 + 2 g_0i k^i k0 + g_ij k^i k^j = 0 with the reference's
 * root choice and fallbacks (MTTN001A.cpp:171-237). near > 0 (re-normalisation in flight):
 * the root nearest the current k^0, since inside the ergoregion (g_00 > 0) both roots can be
 * positive and the first-positive-root rule can take the wrong branch. */
static double null_k0(const double g[16], const double k[4], double near) {
    double A = g[0], B = 0, C = 0;
    for (int i = 1; i < 4; ++i) B += 2.0 * g[i] * k[i];
    for (int i = 1; i < 4; ++i) for (int j = 1; j < 4; ++j) C += g[i * 4 + j] * k[i] * k[j];
    if (fabs(A) < 1e-15) return fabs(B) > 1e-15 ? -C / B : 1.0;
    double disc = B * B - 4.0 * A * C;
    if (disc < 0) return 1.0;
    double s = sqrt(disc);
    double kp = (-B + s) / (2.0 * A), km = (-B - s) / (2.0 * A);
    if (near > 0) return fabs(kp - near) < fabs(km - near) ? kp : km;
    if (kp > 0) return kp;
    if (km > 0) return km;
    return kp > km ? kp : km;
}

static void lower(const double g[16], const double k[4], double p[4]) {
    for (int mu = 0; mu < 4; ++mu) {
        double s = 0;
        for (int nu = 0; nu < 4; ++nu) s += g[mu * 4 + nu] * k[nu];
        p[mu] = s;
    }
}

/* Wormhole: orbital-plane frame of a camera ray. e1 = radial direction of the observer,
 * e2 = unit tangential part of the ray direction. */
typedef struct { double e1[3], e2[3]; } plane_frame;

static void sph_basis(double th, double ph, double er[3], double et[3], double ep[3]) {
    double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
    er[0] = st * cp; er[1] = st * sp; er[2] = ct;
    et[0] = ct * cp; et[1] = ct * sp; et[2] = -st;
    ep[0] = -sp; ep[1] = cp; ep[2] = 0;
}

static void camera_cartesian(const SiriusHipCamera* c, int px, int py, double pos[3], double n[3],
                             double er_out[3]) {
    double th = c->theta_obs;
    if (th < 0.01) th = 0.01;
    if (th > M_PI - 0.01) th = M_PI - 0.01;
    double er[3], et[3], ep[3], ns[3];
    sph_basis(th, c->phi_obs, er, et, ep);
    camera_direction(c, px, py, ns);
    for (int i = 0; i < 3; ++i) {
        pos[i] = c->r_obs * er[i];
        n[i] = ns[0] * er[i] + ns[1] * et[i] + ns[2] * ep[i];
        if (er_out) er_out[i] = er[i];
    }
}

static void mt_frame(const SiriusHipCamera* c, int px, int py, plane_frame* f, double* nr,
                     double* nt) {
    double pos[3], n[3], er[3];
    camera_cartesian(c, px, py, pos, n, er);
    double d = n[0] * er[0] + n[1] * er[1] + n[2] * er[2];
    double t[3] = {n[0] - d * er[0], n[1] - d * er[1], n[2] - d * er[2]};
    double tl = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    if (tl < 1e-300) {
        double et[3], ep[3], e1[3];
        sph_basis(c->theta_obs, c->phi_obs, e1, et, ep);
        t[0] = et[0]; t[1] = et[1]; t[2] = et[2]; tl = 1.0;
    }
    for (int i = 0; i < 3; ++i) { f->e1[i] = er[i]; f->e2[i] = t[i] / tl; }
    *nr = d;
    *nt = tl < 1e-300 ? 0.0 : tl;
}

void oracle_camera_state(const SiriusHipCamera* c, const SiriusHipMetric* m, int px, int py,
                         double x[4], double p[4]) {
    double g[16], dg[64], k[4];
    if (m->family == SIRIUS_METRIC_MORRIS_THORNE) {
        plane_frame f;
        double nr, nt;
        mt_frame(c, px, py, &f, &nr, &nt);
        double b0 = m->throat_b0;
        double l = sqrt(fmax(c->r_obs * c->r_obs - b0 * b0, 0.0));
        double r = c->r_obs;
        x[0] = 0; x[1] = l; x[2] = 0; x[3] = 0;
        /* unit speed in the static orthonormal frame: k^l = n_r, r dpsi/dlam = n_t, k^t = 1 */
        k[0] = 1.0; k[1] = nr; k[2] = nt / r; k[3] = 0;
        oracle_metric(m, x, g, NULL, dg);
        lower(g, k, p);
        return;
    }
    double pos[3], n[3];
    camera_cartesian(c, px, py, pos, n, NULL);
    x[0] = 0; x[1] = pos[0]; x[2] = pos[1]; x[3] = pos[2];
    k[0] = 0; k[1] = n[0]; k[2] = n[1]; k[3] = n[2];
    oracle_metric(m, x, g, NULL, dg);
    k[0] = null_k0(g, k, 0.0);
    lower(g, k, p);
}

/* ========================================================================================
 * Critical curve (Bardeen 1973) and a ray's distance from it
 * ======================================================================================== */
#define CC_SAMPLES 4096
#define CC_TABLE 512

/* Radial table of the critical curve seen from polar angle theta_o: rho_c(psi) on psi in
 * [0, pi] (uniform, CC_TABLE samples) about the centre (alpha0, 0), where alpha0 is the
 * midpoint of the curve's alpha range. Spherical photon orbits of Kerr-Newman: with
 * D = r^2 - 2Mr + a^2 + Q^2, xi(r) = ((r^2 + a^2) D' - 4 r D) / (a D'),
 * eta = (r^2 + a^2 - a xi)^2 / D - (xi - a)^2, alpha = -xi / sin th, beta^2 = eta +
 * a^2 cos^2 th - xi^2 cot^2 th. A non-rotating hole gives a circle of radius
 * r_ph^2 / sqrt(D(r_ph)). */
static void crit_table(const SiriusHipMetric* m, double theta_o, double* tab, double* alpha0) {
    double M = m->M, a = m->a, Q2 = m->Q * m->Q;
    if (fabs(a) < 1e-6) {
        double rph = 0.5 * (3 * M + sqrt(fmax(9 * M * M - 8 * Q2, 0)));
        double D = rph * rph - 2 * M * rph + Q2;
        for (int j = 0; j < CC_TABLE; ++j) tab[j] = rph * rph / sqrt(D);
        *alpha0 = 0.0;
        return;
    }
    static __thread double al[CC_SAMPLES], be[CC_SAMPLES], ps[CC_SAMPLES], rh[CC_SAMPLES];
    double rp = oracle_horizon(m);
    double st = sin(theta_o), ct = cos(theta_o);
    int n = 0;
    double amin = 1e300, amax = -1e300;
    for (int i = 0; i < CC_SAMPLES; ++i) {
        double r = rp + (5.0 * M - rp) * (i + 0.5) / CC_SAMPLES;
        double D = r * r - 2 * M * r + a * a + Q2, Dp = 2 * r - 2 * M;
        if (D <= 0 || fabs(Dp) < 1e-12) continue;
        double xi = ((r * r + a * a) * Dp - 4 * r * D) / (a * Dp);
        double t1 = r * r + a * a - a * xi;
        double eta = t1 * t1 / D - (xi - a) * (xi - a);
        if (eta < 0) continue;
        double b2 = eta + a * a * ct * ct - xi * xi * ct * ct / (st * st);
        if (b2 < 0) continue;
        al[n] = -xi / st;
        be[n] = sqrt(b2);
        if (al[n] < amin) amin = al[n];
        if (al[n] > amax) amax = al[n];
        ++n;
    }
    double c0 = 0.5 * (amin + amax);
    *alpha0 = c0;
    for (int i = 0; i < n; ++i) { ps[i] = atan2(be[i], al[i] - c0); rh[i] = hypot(al[i] - c0, be[i]); }
    /* insertion sort by psi (the samples are already nearly ordered) */
    for (int i = 1; i < n; ++i) {
        double p0 = ps[i], r0 = rh[i];
        int j = i - 1;
        while (j >= 0 && ps[j] > p0) { ps[j + 1] = ps[j]; rh[j + 1] = rh[j]; --j; }
        ps[j + 1] = p0; rh[j + 1] = r0;
    }
    int k = 0;
    for (int j = 0; j < CC_TABLE; ++j) {
        double psi = M_PI * j / (CC_TABLE - 1);
        if (n == 0) { tab[j] = 0; continue; }
        if (psi <= ps[0]) { tab[j] = rh[0]; continue; }
        if (psi >= ps[n - 1]) { tab[j] = rh[n - 1]; continue; }
        while (k < n - 2 && ps[k + 1] < psi) ++k;
        double w = (psi - ps[k]) / (ps[k + 1] - ps[k]);
        tab[j] = rh[k] + w * (rh[k + 1] - rh[k]);
    }
}

/* (alpha, beta) of a ray from its constants of motion: alpha = -L / (E sin th),
 * beta = p_th / E at the ray's start (exact for the theta potential of Kerr-Newman). */
static void ray_alpha_beta(const SiriusHipMetric* m, const double x[4], const double p[4],
                           double* al, double* be, double* theta) {
    double E = -p[0];
    double L = x[1] * p[2] - x[2] * p[1];
    double r = oracle_kerr_radius(m, x[1], x[2], x[3]);
    double ct = x[3] / r, st = sqrt(fmax(1 - ct * ct, 1e-300));
    double pth = (ct / st) * (x[1] * p[1] + x[2] * p[2]) - r * st * p[3];
    *al = -L / (E * st);
    *be = pth / E;
    *theta = acos(ct);
}

static double crit_lookup(const double* tab, double alpha0, double al, double be) {
    double da = al - alpha0, b = fabs(be);
    double psi = atan2(b, da), rho = hypot(da, b);
    double u = psi / M_PI * (CC_TABLE - 1);
    int j = (int)u;
    if (j < 0) j = 0;
    if (j > CC_TABLE - 2) j = CC_TABLE - 2;
    double w = u - j;
    return rho - (tab[j] + w * (tab[j + 1] - tab[j]));
}

double oracle_critical_distance(const SiriusHipCamera* c, const SiriusHipMetric* m, int px, int py) {
    if (m->family != SIRIUS_METRIC_KERR_SCHILD) return 1e30;
    double x[4], p[4], al, be, th, tab[CC_TABLE], a0;
    oracle_camera_state(c, m, px, py, x, p);
    ray_alpha_beta(m, x, p, &al, &be, &th);
    crit_table(m, th, tab, &a0);
    return crit_lookup(tab, a0, al, be);
}

/* ========================================================================================
 * Dormand-Prince 5(4) (PHGD001A.cpp:16-79) over y = (x^mu, p_mu)
 * ======================================================================================== */
static const double A21 = 1.0 / 5.0;
static const double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
static const double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
static const double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                    A54 = -212.0 / 729.0;
static const double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                    A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
static const double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                    B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
static const double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0,
                    E5 = -17253.0 / 339200.0, E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;

typedef struct { double y[8]; } st8;      /* (x^0..3, p_0..3) */
typedef struct { double f[8]; double pk; } der8; /* (dx, dp) and p.k */

static void eval_der(const SiriusHipMetric* m, const st8* s, der8* d) {
    d->pk = oracle_rhs(m, s->y, s->y + 4, d->f, d->f + 4);
}

/* One DP45 step of size h from (s0, d0). Writes the 5th-order solution, its derivative
 * (FSAL) and, when err != NULL, the position error vector err[4]. */
static void dp45(const SiriusHipMetric* m, const st8* s0, const der8* d0, double h, st8* s1,
                 der8* d1, double err[4]) {
    der8 k2, k3, k4, k5, k6;
    st8 t;
    const double* k1 = d0->f;
    for (int i = 0; i < 8; ++i) t.y[i] = s0->y[i] + h * A21 * k1[i];
    eval_der(m, &t, &k2);
    for (int i = 0; i < 8; ++i) t.y[i] = s0->y[i] + h * (A31 * k1[i] + A32 * k2.f[i]);
    eval_der(m, &t, &k3);
    for (int i = 0; i < 8; ++i) t.y[i] = s0->y[i] + h * (A41 * k1[i] + A42 * k2.f[i] + A43 * k3.f[i]);
    eval_der(m, &t, &k4);
    for (int i = 0; i < 8; ++i)
        t.y[i] = s0->y[i] + h * (A51 * k1[i] + A52 * k2.f[i] + A53 * k3.f[i] + A54 * k4.f[i]);
    eval_der(m, &t, &k5);
    for (int i = 0; i < 8; ++i)
        t.y[i] = s0->y[i] + h * (A61 * k1[i] + A62 * k2.f[i] + A63 * k3.f[i] + A64 * k4.f[i] +
                                 A65 * k5.f[i]);
    eval_der(m, &t, &k6);
    for (int i = 0; i < 8; ++i)
        s1->y[i] = s0->y[i] + h * (B1 * k1[i] + B3 * k3.f[i] + B4 * k4.f[i] + B5 * k5.f[i] +
                                   B6 * k6.f[i]);
    eval_der(m, s1, d1);
    if (err)
        for (int i = 0; i < 4; ++i)
            err[i] = h * (E1 * k1[i] + E3 * k3.f[i] + E4 * k4.f[i] + E5 * k5.f[i] + E6 * k6.f[i] +
                          E7 * d1->f[i]);
}

/* computeOptimalStep (PHGD001A.cpp:471-483) with the same constants; hmax is the step cap
 * (h_max, or the far-field cap max(h_max, far_slope r) of the FP32 contract). */
static double next_step_cap(double h, double err, double hmin, double hmax) {
    double hn;
    if (err < 1e-15) hn = h * 2.0;
    else {
        double fac = 0.9 * pow(1.0 / err, 0.2);
        if (fac < 0.1) fac = 0.1;
        if (fac > 2.0) fac = 2.0;
        hn = h * fac;
    }
    if (hn < hmin) hn = hmin;
    if (hn > hmax) hn = hmax;
    return hn;
}

static double step_cap(const SiriusHipMetric* m, const SiriusHipTraceOpts* o, const double x[4]) {
    if (o->far_slope <= 0) return o->h_max;
    double r = m->family == SIRIUS_METRIC_MORRIS_THORNE ? fabs(x[1]) : oracle_kerr_radius(m, x[1], x[2], x[3]);
    double c = o->far_slope * r;
    return c > o->h_max ? c : o->h_max;
}

/* Cubic Hermite interpolant of the position over a step, theta in [0, 1]. */
static void hermite(const st8* s0, const der8* d0, const st8* s1, const der8* d1, double h,
                    double th, double out[4]) {
    double t2 = th * th, t3 = t2 * th;
    double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + th, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
    for (int i = 0; i < 4; ++i)
        out[i] = h00 * s0->y[i] + h10 * h * d0->f[i] + h01 * s1->y[i] + h11 * h * d1->f[i];
}

enum { EV_Z = 0, EV_R = 1, EV_L = 2 };

static double event_value(const SiriusHipMetric* m, int kind, const double x[4], double target) {
    if (kind == EV_Z) return x[3];
    if (kind == EV_L) return fabs(x[1]) - target;
    return oracle_kerr_radius(m, x[1], x[2], x[3]) - target;
}

/* Bisection on the Hermite interpolant for the event's sign change inside the step. */
static double locate_event(const SiriusHipMetric* m, int kind, double target, const st8* s0,
                           const der8* d0, const st8* s1, const der8* d1, double h) {
    double lo = 0, hi = 1, x[4];
    double flo = event_value(m, kind, s0->y, target);
    for (int it = 0; it < 60; ++it) {
        double mid = 0.5 * (lo + hi);
        hermite(s0, d0, s1, d1, h, mid, x);
        double fm = event_value(m, kind, x, target);
        if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; } else hi = mid;
    }
    return 0.5 * (lo + hi);
}

/* Land on the event: one DP45 step of size th*h from s0, then one Newton correction of th
 * from the event's rate at the landing point. Returns the final fraction. */
static double land(const SiriusHipMetric* m, int kind, double target, const st8* s0,
                   const der8* d0, double h, double th, st8* sl, der8* dl, int newton) {
    dp45(m, s0, d0, th * h, sl, dl, NULL);
    if (newton) {
        double v = event_value(m, kind, sl->y, target), rate;
        if (kind == EV_Z) rate = dl->f[3];
        else {
            double eps = 1e-6;
            double xp[4];
            for (int i = 0; i < 4; ++i) xp[i] = sl->y[i] + eps * dl->f[i];
            rate = (event_value(m, kind, xp, target) - v) / eps;
        }
        if (fabs(rate) > 1e-300) {
            double th2 = th - v / (rate * h);
            if (th2 > 0 && th2 < 1.5) {
                th = th2;
                dp45(m, s0, d0, th * h, sl, dl, NULL);
            }
        }
    }
    return th;
}

static int is_bad(const st8* s) {
    for (int i = 0; i < 8; ++i) if (!isfinite(s->y[i])) return 1;
    return 0;
}

/* Per-camera constants shared by all its rays. */
typedef struct {
    const SiriusHipMetric* m;
    double r_cap, r_in, r_out, fmax, obs_norm;
    double bins[64];
} cam_ctx;

static void setup_cam(const SiriusHipCamera* c, const SiriusHipMetric* m,
                      const SiriusHipTraceOpts* o, cam_ctx* cx) {
    cx->m = m;
    cx->r_cap = o->horizon_factor * oracle_horizon(m);
    cx->r_in = c->disk_r_in > 0 ? c->disk_r_in : oracle_isco(m->a);
    cx->r_out = c->disk_r_out;
    cx->fmax = (m->family == SIRIUS_METRIC_KERR_SCHILD) ? oracle_disk_tmax_f_over_r(m->a, cx->r_in, cx->r_out) : 0;
    oracle_bin_centres(o->n_bins, o->x_min, o->x_max, cx->bins);
    /* static observer at the camera: u = d_t / sqrt(-g_tt) */
    double x[4], p[4], g[16], dg[64];
    oracle_camera_state(c, m, c->width / 2, c->height / 2, x, p);
    oracle_metric(m, x, g, NULL, dg);
    cx->obs_norm = 1.0 / sqrt(-g[0]);
}

/* Thin-disk emission at the landed hit: exact g = (p.u_obs)/(p.u_emit) with a Keplerian
 * (Kerr-Newman, prograde) emitter and a static observer; I_k = x_k^3 / (e^{x_k/(g tau)} - 1). */
static void disk_emission(const cam_ctx* cx, const SiriusHipTraceOpts* o, const st8* s, double r,
                          double* gout, double* I) {
    const SiriusHipMetric* m = cx->m;
    double M = m->M, a = m->a, Q2 = m->Q * m->Q;
    double sq = sqrt(fmax(M * r - Q2, 0.0));
    double Om = sq / (r * r + a * sq);
    double g[16], dg[64];
    oracle_metric(m, s->y, g, NULL, dg);
    double v[4] = {1.0, -Om * s->y[2], Om * s->y[1], 0.0};
    double vv = 0;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) vv += g[i * 4 + j] * v[i] * v[j];
    double ut = 1.0 / sqrt(-vv);
    const double* p = s->y + 4;
    double pphi = s->y[1] * p[2] - s->y[2] * p[1];
    double gf = (p[0] * cx->obs_norm) / (ut * (p[0] + Om * pphi));
    double f = oracle_page_thorne_f(r, a) / r;
    double tau = (cx->fmax > 0 && f > 0) ? pow(f / cx->fmax, 0.25) : 0.0;
    *gout = gf;
    for (int k = 0; k < o->n_bins; ++k) {
        double xk = cx->bins[k];
        I[k] = (tau > 0 && gf > 0) ? xk * xk * xk / expm1(xk / (gf * tau)) : 0.0;
    }
}

/* Cubic Hermite interpolant of all 8 state components (p uses dp/dlam). */
static void hermite8(const st8* s0, const der8* d0, const st8* s1, const der8* d1, double h,
                     double th, double out[8]) {
    double t2 = th * th, t3 = t2 * th;
    double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + th, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
    for (int i = 0; i < 8; ++i)
        out[i] = h00 * s0->y[i] + h10 * h * d0->f[i] + h01 * s1->y[i] + h11 * h * d1->f[i];
}

/* Step cap inside the emitting layer: a step may travel at most half-way to the 3-sigma
 * layer and never more than 0.2 sigma inside it (sigma = h_over_r * r). */
static double vol_step_cap(const SiriusHipMetric* m, const SiriusHipTraceOpts* o, const double x[4]) {
    double r = oracle_kerr_radius(m, x[1], x[2], x[3]);
    if (r < 0.5 * o->vol_r_min || r > 1.5 * o->vol_r_max) return 1e30;
    double sig = o->vol_h_over_r * r;
    double c = 0.5 * (fabs(x[3]) - 3.0 * sig);
    return c > 0.2 * sig ? c : 0.2 * sig;
}

/* Smooth (C2) radial window of the emitting shell: 0 at r_min and r_max, 1 on
 * [1.1 r_min, 0.9 r_max], quintic smoothstep 6u^5 - 15u^4 + 10u^3 between. */
static double vol_window(double r, double r_min, double r_max) {
    double u = (r - r_min) / (0.1 * r_min), v = (r_max - r) / (0.1 * r_max);
    u = u < 0 ? 0 : (u > 1 ? 1 : u);
    v = v < 0 ? 0 : (v > 1 ? 1 : v);
    return u * u * u * (u * (6 * u - 15) + 10) * v * v * v * (v * (6 * v - 15) + 10);
}

/* Volumetric transfer over one accepted step, 3-node Gauss-Legendre on the Hermite
 * interpolant. Invariant form (I/nu^3, j/nu^2, nu alpha): with g = nu_obs/nu_emit,
 *   dtau = kappa rho nu_e dlam,  dI_k = rho S(x_k / g) (nu_o^3 / nu_e^2) e^{-tau} dlam,
 * rho = j0 r^-2 exp(-z^2/(2 sigma^2)) W(r) (W: the shell window above, from max(r_min, r_in) to
 * r_max: inside the ISCO (r_in by default) a Keplerian emitter is unstable and, nearer the
 * photon orbit, ultra-relativistic, so the emitting shell starts there),
 * S(x) = x^3/(e^{x/tau_v} - 1), tau_v = (6/r)^{3/4},
 * emitter on a Keplerian (Kerr-Newman) circle at r_c = sqrt(x^2 + y^2 - a^2). This is the
 * build's definition of the volumetric case; the reference's gray straight-chord march
 * (GTRC001A.cpp:914-998) is the role it replaces. */
static void vol_accumulate(const cam_ctx* cx, const SiriusHipTraceOpts* o, const st8* s0,
                           const der8* d0, const st8* s1, const der8* d1, double h, double* I,
                           double* tau_acc) {
    static const double gn[3] = {0.1127016653792583, 0.5, 0.8872983346207417};
    static const double gw[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    const SiriusHipMetric* m = cx->m;
    for (int q = 0; q < 3; ++q) {
        double y8[8];
        hermite8(s0, d0, s1, d1, h, gn[q], y8);
        const double* x = y8;
        const double* p = y8 + 4;
        double r = oracle_kerr_radius(m, x[1], x[2], x[3]);
        const double r_lo = cx->r_in > o->vol_r_min ? cx->r_in : o->vol_r_min; /* no emitters inside the ISCO */
        if (r < r_lo || r > o->vol_r_max) continue;
        double sig = o->vol_h_over_r * r;
        double rho = o->vol_j0 / (r * r) * exp(-x[3] * x[3] / (2 * sig * sig)) *
                     vol_window(r, r_lo, o->vol_r_max);
        double M = m->M, a = m->a, Q2 = m->Q * m->Q;
        double rc = sqrt(fmax(x[1] * x[1] + x[2] * x[2] - a * a, 1e-12));
        double sq = sqrt(fmax(M * rc - Q2, 0.0));
        double Om = sq / (rc * rc + a * sq);
        double g[16], dg[64];
        oracle_metric(m, x, g, NULL, dg);
        double v[4] = {1.0, -Om * x[2], Om * x[1], 0.0};
        double vv = 0;
        for (int i = 0; i < 4; ++i) for (int k = 0; k < 4; ++k) vv += g[i * 4 + k] * v[i] * v[k];
        if (vv >= 0) continue;
        double ut = 1.0 / sqrt(-vv);
        double nue = -ut * (p[0] + Om * (x[1] * p[2] - x[2] * p[1]));
        double nuo = -p[0] * cx->obs_norm;
        if (nue <= 0 || nuo <= 0) continue;
        double gf = nuo / nue;
        double dl = gw[q] * h;
        double dtau = o->vol_kappa * rho * nue * dl;
        double w = rho * (nuo * nuo * nuo / (nue * nue)) * dl * exp(-(*tau_acc + 0.5 * dtau));
        double tv = pow(6.0 / r, 0.75);
        for (int k = 0; k < o->n_bins; ++k) {
            double xe = cx->bins[k] / gf;
            I[k] += w * xe * xe * xe / expm1(xe / tv);
        }
        *tau_acc += dtau;
    }
}

typedef struct {
    int cls, flags;
    double theta, phi, lambda, disk_r, g, null, I[64];
    uint32_t steps, rejected, landings;
    double state[8];
} ray_out;

/* The trace contract for one ray (see the header comment for the departures). */
static void trace_dp45(const cam_ctx* cx, const SiriusHipTraceOpts* o, const st8* init, ray_out* R) {
    const SiriusHipMetric* m = cx->m;
    int fam = m->family;
    st8 s0 = *init, s1;
    der8 d0, d1;
    eval_der(m, &s0, &d0);
    double h = o->h_init, lam = 0, err4[4];
    /* with a far-field cap the first step starts at the cap (the FP32 path's policy); the
     * reference's h0 = 0.01 would spend ~10 doubling steps reaching it from the observer */
    if (o->far_slope > 0) {
        double cap = step_cap(m, o, s0.y);
        if (cap > h) h = cap;
    }
    double tau = 0;
    memset(R, 0, sizeof(*R));
    R->cls = SIRIUS_RAY_MAX_STEPS;
    int attempts = 0;
    int vol = (fam == SIRIUS_METRIC_KERR_SCHILD && o->emission == SIRIUS_EMISSION_VOLUMETRIC);
    while (attempts < o->max_steps) {
        if (vol) {
            double cap = vol_step_cap(m, o, s0.y);
            if (h > cap) h = cap;
        }
        dp45(m, &s0, &d0, h, &s1, &d1, err4);
        attempts++;
        double e = 0;
        for (int i = 0; i < 4; ++i) {
            double sc = o->atol + o->rtol * fabs(s1.y[i]);
            double q = err4[i] / sc;
            e += q * q;
        }
        e = sqrt(e / 4.0);
        if (!(e <= 1.0)) {
            if (!isfinite(e)) { R->cls = SIRIUS_RAY_FAILED; break; }
            h = next_step_cap(h, e, o->h_min, step_cap(m, o, s0.y));
            R->rejected++;
            if (h <= o->h_min) { R->cls = SIRIUS_RAY_FAILED; break; }
            continue;
        }
        if (is_bad(&s1)) { R->cls = SIRIUS_RAY_FAILED; break; }
        /* ---- events inside the accepted step ---- */
        if (vol) {
            vol_accumulate(cx, o, &s0, &d0, &s1, &d1, h, R->I, &tau);
            if (tau > o->vol_tau_max) {
                lam += h;
                R->cls = SIRIUS_RAY_DISK;
                break;
            }
        }
        if (fam == SIRIUS_METRIC_KERR_SCHILD && o->emission == SIRIUS_EMISSION_THIN_DISK &&
            s0.y[3] * s1.y[3] < 0) {
            /* The hit is decided on the Hermite-interpolated crossing radius; the reported
             * radius and the emission come from the landed state. */
            double th = locate_event(m, EV_Z, 0, &s0, &d0, &s1, &d1, h);
            double xc[4];
            hermite(&s0, &d0, &s1, &d1, h, th, xc);
            double rc = oracle_kerr_radius(m, xc[1], xc[2], xc[3]);
            if (rc >= cx->r_in && rc <= cx->r_out) {
                st8 sl; der8 dl;
                th = land(m, EV_Z, 0, &s0, &d0, h, th, &sl, &dl, 1);
                R->landings += 2;
                double rd = oracle_kerr_radius(m, sl.y[1], sl.y[2], sl.y[3]);
                R->cls = SIRIUS_RAY_DISK;
                R->disk_r = rd;
                disk_emission(cx, o, &sl, rd, &R->g, R->I);
                lam += th * h;
                s1 = sl; d1 = dl;
                break;
            }
        }
        double r1;
        if (fam == SIRIUS_METRIC_MORRIS_THORNE) r1 = fabs(s1.y[1]);
        else r1 = oracle_kerr_radius(m, s1.y[1], s1.y[2], s1.y[3]);
        if (fam == SIRIUS_METRIC_KERR_SCHILD && r1 < cx->r_cap) {
            double r0 = oracle_kerr_radius(m, s0.y[1], s0.y[2], s0.y[3]);
            double th = 1.0;
            if (r0 > cx->r_cap) {
                th = locate_event(m, EV_R, cx->r_cap, &s0, &d0, &s1, &d1, h);
                st8 sl; der8 dl;
                th = land(m, EV_R, cx->r_cap, &s0, &d0, h, th, &sl, &dl, 0);
                R->landings++;
                s1 = sl; d1 = dl;
            }
            lam += th * h;
            R->cls = SIRIUS_RAY_CAPTURED;
            break;
        }
        if (r1 > o->r_escape) {
            double radial;
            if (fam == SIRIUS_METRIC_MORRIS_THORNE) radial = s1.y[1] * d1.f[1];
            else radial = s1.y[1] * d1.f[1] + s1.y[2] * d1.f[2] + s1.y[3] * d1.f[3];
            if (radial > 0) {
                double r0 = (fam == SIRIUS_METRIC_MORRIS_THORNE) ? fabs(s0.y[1])
                           : oracle_kerr_radius(m, s0.y[1], s0.y[2], s0.y[3]);
                double th = 1.0;
                if (r0 < o->r_escape) {
                    int kind = fam == SIRIUS_METRIC_MORRIS_THORNE ? EV_L : EV_R;
                    th = locate_event(m, kind, o->r_escape, &s0, &d0, &s1, &d1, h);
                    st8 sl; der8 dl;
                    th = land(m, kind, o->r_escape, &s0, &d0, h, th, &sl, &dl, 0);
                    R->landings++;
                    s1 = sl; d1 = dl;
                }
                lam += th * h;
                R->cls = SIRIUS_RAY_ESCAPED;
                break;
            }
        }
        /* ---- accept ---- */
        lam += h;
        {
            /* inward steps scale the proposal by (r1 / r0)^step_radial_exp (0: the reference's
             * controller, computeOptimalStep, PHGD001A.cpp:471-483) */
            double hp = h;
            if (o->step_radial_exp > 0) {
                double r0 = fam == SIRIUS_METRIC_MORRIS_THORNE ? fabs(s0.y[1])
                            : oracle_kerr_radius(m, s0.y[1], s0.y[2], s0.y[3]);
                if (r1 < r0) hp *= pow(r1 / r0, o->step_radial_exp);
            }
            h = next_step_cap(hp, e, o->h_min, step_cap(m, o, s1.y));
        }
        s0 = s1;
        d0 = d1;
        /* null re-normalisation when |g k k| > 1e-6 (PHGD001A.cpp:609-638) */
        if (fabs(d0.pk) > 1e-6) {
            double g[16], dg[64], k[4];
            oracle_metric(m, s0.y, g, NULL, dg);
            for (int i = 0; i < 4; ++i) k[i] = d0.f[i];
            double k0 = null_k0(g, k, k[0]);
            if (isfinite(k0)) {
                k[0] = k0;
                lower(g, k, s0.y + 4);
                eval_der(m, &s0, &d0);
                R->flags |= SIRIUS_FLAG_RENORM;
            }
        }
    }
    R->steps = (uint32_t)attempts;
    R->lambda = lam;
    R->null = fabs(d1.pk);
    if (R->cls == SIRIUS_RAY_MAX_STEPS || R->cls == SIRIUS_RAY_FAILED) {
        R->null = fabs(d0.pk);
        memcpy(R->state, s0.y, sizeof(R->state));
    } else {
        memcpy(R->state, s1.y, sizeof(R->state));
    }
    if (R->cls == SIRIUS_RAY_ESCAPED) {
        double dir[3];
        if (fam == SIRIUS_METRIC_MORRIS_THORNE) {
            /* handled by the caller (needs the camera frame) */
            dir[0] = d1.f[1]; dir[1] = d1.f[2]; dir[2] = 0;
        } else {
            dir[0] = d1.f[1]; dir[1] = d1.f[2]; dir[2] = d1.f[3];
        }
        double n = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        R->theta = acos(fmax(-1.0, fmin(1.0, dir[2] / n)));
        R->phi = atan2(dir[1], dir[0]);
    }
    if (vol) R->g = tau;  /* optical depth travelled, reported in the redshift slot */
}

/* Wormhole: map the planar end state back to a sky direction on the exit side. */
static void mt_direction(const SiriusHipCamera* c, const SiriusHipMetric* m, int px, int py,
                         ray_out* R) {
    plane_frame f;
    double nr, nt;
    mt_frame(c, px, py, &f, &nr, &nt);
    double l = R->state[1], psi = R->state[2], pl = R->state[5], L = R->state[6];
    double b0 = m->throat_b0, r = sqrt(l * l + b0 * b0);
    double drdl = (l / r) * pl, rdpsi = L / r;
    double cp = cos(psi), sp = sin(psi), e[3], t[3], d[3];
    for (int i = 0; i < 3; ++i) {
        e[i] = cp * f.e1[i] + sp * f.e2[i];
        t[i] = -sp * f.e1[i] + cp * f.e2[i];
        d[i] = drdl * e[i] + rdpsi * t[i];
    }
    double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    R->theta = acos(fmax(-1.0, fmin(1.0, d[2] / n)));
    R->phi = atan2(d[1], d[0]);
    if (l < 0) R->flags |= SIRIUS_FLAG_FAR_SIDE;
}

/* ========================================================================================
 * Symplectic integrator: Tao (2016) extended phase space, Strang split, Yoshida composition
 * (weights as PHSI001A.h:43-80). The reference applies explicit leapfrog to the
 * non-separable H (PHSI001A.h:430-455); Tao's doubling is explicit and symplectic in the
 * doubled space, so it is used instead and pinned by H, E and L_z drift.
 * ======================================================================================== */
static void tao_flow_a(const SiriusHipMetric* m, double d, double q[4], double p[4], double x[4],
                       double y[4]) {
    double dq[4], dp[4];
    oracle_rhs(m, q, y, dq, dp); /* H(q, y) */
    for (int i = 0; i < 4; ++i) { p[i] += d * dp[i]; x[i] += d * dq[i]; }
}
static void tao_flow_b(const SiriusHipMetric* m, double d, double q[4], double p[4], double x[4],
                       double y[4]) {
    double dq[4], dp[4];
    oracle_rhs(m, x, p, dq, dp); /* H(x, p) */
    for (int i = 0; i < 4; ++i) { q[i] += d * dq[i]; y[i] += d * dp[i]; }
}
static void tao_flow_c(double d, double w, double q[4], double p[4], double x[4], double y[4]) {
    double c = cos(2 * w * d), s = sin(2 * w * d);
    for (int i = 0; i < 4; ++i) {
        double sq = q[i] + x[i], sp = p[i] + y[i], dq = q[i] - x[i], dp = p[i] - y[i];
        double ndq = c * dq + s * dp, ndp = -s * dq + c * dp;
        q[i] = 0.5 * (sq + ndq); p[i] = 0.5 * (sp + ndp);
        x[i] = 0.5 * (sq - ndq); y[i] = 0.5 * (sp - ndp);
    }
}
static void tao_step2(const SiriusHipMetric* m, double h, double w, double q[4], double p[4],
                      double x[4], double y[4]) {
    tao_flow_a(m, 0.5 * h, q, p, x, y);
    tao_flow_b(m, 0.5 * h, q, p, x, y);
    tao_flow_c(h, w, q, p, x, y);
    tao_flow_b(m, 0.5 * h, q, p, x, y);
    tao_flow_a(m, 0.5 * h, q, p, x, y);
}
static const double YO4[3] = {1.3512071919596576, -1.7024143839193153, 1.3512071919596576};

static void trace_symplectic(const cam_ctx* cx, const SiriusHipTraceOpts* o, const st8* init,
                             ray_out* R) {
    const SiriusHipMetric* m = cx->m;
    double q[4], p[4], x[4], y[4];
    for (int i = 0; i < 4; ++i) { q[i] = x[i] = init->y[i]; p[i] = y[i] = init->y[4 + i]; }
    double h = o->symplectic_h, w = 20.0, lam = 0;
    memset(R, 0, sizeof(*R));
    R->cls = SIRIUS_RAY_MAX_STEPS;
    int n = 0;
    double rprev = oracle_kerr_radius(m, q[1], q[2], q[3]);
    double zprev = q[3];
    while (n < o->max_steps) {
        double q0[4], p0[4];
        memcpy(q0, q, sizeof(q0)); memcpy(p0, p, sizeof(p0));
        if (o->symplectic_order >= 4)
            for (int s = 0; s < 3; ++s) tao_step2(m, YO4[s] * h, w, q, p, x, y);
        else tao_step2(m, h, w, q, p, x, y);
        n++;
        lam += h;
        int bad = 0;
        for (int i = 0; i < 4; ++i) if (!isfinite(q[i]) || !isfinite(p[i])) bad = 1;
        if (bad) { R->cls = SIRIUS_RAY_FAILED; break; }
        double r = oracle_kerr_radius(m, q[1], q[2], q[3]);
        if (m->family == SIRIUS_METRIC_KERR_SCHILD && o->emission == SIRIUS_EMISSION_THIN_DISK &&
            zprev * q[3] < 0) {
            /* the crossing is located on the cubic Hermite interpolant of the whole state over
             * the step (the derivatives of H at both ends), and the ray ends there: a linear
             * interpolant would leave an O(h^2) null-constraint error of its own at the hit */
            st8 s0, s1, sc;
            der8 d0, d1;
            for (int i = 0; i < 4; ++i) { s0.y[i] = q0[i]; s0.y[4 + i] = p0[i]; s1.y[i] = q[i]; s1.y[4 + i] = p[i]; }
            eval_der(m, &s0, &d0);
            eval_der(m, &s1, &d1);
            double th = locate_event(m, EV_Z, 0, &s0, &d0, &s1, &d1, h);
            hermite8(&s0, &d0, &s1, &d1, h, th, sc.y);
            double rd = sqrt(fmax(sc.y[1] * sc.y[1] + sc.y[2] * sc.y[2] - m->a * m->a, 0.0));
            if (rd >= cx->r_in && rd <= cx->r_out) {
                for (int i = 0; i < 4; ++i) { q[i] = sc.y[i]; p[i] = sc.y[4 + i]; }
                lam += (th - 1) * h;
                R->cls = SIRIUS_RAY_DISK;
                R->disk_r = rd;
                disk_emission(cx, o, &sc, rd, &R->g, R->I);
                break;
            }
        }
        if (m->family == SIRIUS_METRIC_KERR_SCHILD && r < cx->r_cap) { R->cls = SIRIUS_RAY_CAPTURED; break; }
        if (r > o->r_escape && r > rprev) { R->cls = SIRIUS_RAY_ESCAPED; break; }
        rprev = r;
        zprev = q[3];
    }
    double dq[4], dp[4];
    double pk = oracle_rhs(m, q, p, dq, dp);
    R->null = fabs(pk);
    R->steps = (uint32_t)n;
    R->lambda = lam;
    for (int i = 0; i < 4; ++i) { R->state[i] = q[i]; R->state[4 + i] = p[i]; }
    if (R->cls == SIRIUS_RAY_ESCAPED) {
        double nn = sqrt(dq[1] * dq[1] + dq[2] * dq[2] + dq[3] * dq[3]);
        R->theta = acos(fmax(-1.0, fmin(1.0, dq[3] / nn)));
        R->phi = atan2(dq[2], dq[1]);
    }
}

/* ========================================================================================
 * Batch driver
 * ======================================================================================== */
int oracle_trace(const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics,
                 int n_metric, const SiriusHipTraceOpts* o, const int64_t* ray_ids, int64_t n_ids,
                 SiriusHipRayOutput* out, int n_threads, SiriusHipStats* stats) {
    if (!cams || n_cam <= 0 || !metrics || !o || !out || o->n_bins > 64 || o->n_bins < 0) return SIRIUS_HIP_E_INVALID;
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (n_cam + 1));
    cam_ctx* cx = (cam_ctx*)malloc(sizeof(cam_ctx) * n_cam);
    off[0] = 0;
    for (int c = 0; c < n_cam; ++c) {
        if (cams[c].metric_index < 0 || cams[c].metric_index >= n_metric) { free(off); free(cx); return SIRIUS_HIP_E_INVALID; }
        off[c + 1] = off[c] + (int64_t)cams[c].width * cams[c].height;
        setup_cam(&cams[c], &metrics[cams[c].metric_index], o, &cx[c]);
    }
    int64_t total = off[n_cam];
    if (!ray_ids && n_ids != total) { free(off); free(cx); return SIRIUS_HIP_E_INVALID; }
    uint64_t acc = 0, rej = 0, land_n = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : acc, rej, land_n)
    for (int64_t i = 0; i < n_ids; ++i) {
        int64_t id = ray_ids ? ray_ids[i] : i;
        if (id < 0 || id >= total) continue;
        int c = 0;
        while (off[c + 1] <= id) ++c;
        int64_t pix = id - off[c];
        int px = (int)(pix % cams[c].width), py = (int)(pix / cams[c].width);
        st8 s;
        oracle_camera_state(&cams[c], cx[c].m, px, py, s.y, s.y + 4);
        ray_out R;
        if (o->integrator == SIRIUS_INTEGRATOR_SYMPLECTIC) trace_symplectic(&cx[c], o, &s, &R);
        else trace_dp45(&cx[c], o, &s, &R);
        if (cx[c].m->family == SIRIUS_METRIC_MORRIS_THORNE && R.cls == SIRIUS_RAY_ESCAPED)
            mt_direction(&cams[c], cx[c].m, px, py, &R);
        acc += R.steps - R.rejected;
        rej += R.rejected;
        land_n += R.landings;
        if (out->cls) out->cls[i] = (uint8_t)R.cls;
        if (out->flags) out->flags[i] = (uint8_t)R.flags;
        if (out->theta) out->theta[i] = (float)R.theta;
        if (out->phi) out->phi[i] = (float)R.phi;
        if (out->lambda) out->lambda[i] = (float)R.lambda;
        if (out->disk_r) out->disk_r[i] = (float)R.disk_r;
        if (out->redshift) out->redshift[i] = (float)R.g;
        if (out->null_drift) out->null_drift[i] = (float)R.null;
        if (out->steps) out->steps[i] = R.steps;
        if (out->intensity)
            for (int k = 0; k < o->n_bins; ++k) out->intensity[(int64_t)k * n_ids + i] = (float)R.I[k];
        if (out->state) memcpy(out->state + 8 * i, R.state, sizeof(R.state));
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->rays = (uint64_t)n_ids;
        stats->steps_accepted = acc;
        stats->steps_rejected = rej;
        stats->steps_landing = land_n;
    }
    free(off);
    free(cx);
    return SIRIUS_HIP_OK;
}

/* sizeof/offsetof of the C-ABI structs as a C compiler lays them out (checked by tests against
 * the ctypes mirror). Order: Metric, Camera, TraceOpts, RayOutput, Stats sizes, then offsets
 * of Camera.width, TraceOpts.n_bins, TraceOpts.vol_tau_max, RayOutput.on_device,
 * Stats.kernel_ms, Stats.null_drift_median. */
int oracle_abi_layout(int64_t* out, int n) {
    int64_t v[] = {
        (int64_t)sizeof(SiriusHipMetric), (int64_t)sizeof(SiriusHipCamera), (int64_t)sizeof(SiriusHipTraceOpts),
        (int64_t)sizeof(SiriusHipRayOutput), (int64_t)sizeof(SiriusHipStats),
        (int64_t)offsetof(SiriusHipCamera, width), (int64_t)offsetof(SiriusHipTraceOpts, n_bins),
        (int64_t)offsetof(SiriusHipTraceOpts, vol_tau_max), (int64_t)offsetof(SiriusHipRayOutput, on_device),
        (int64_t)offsetof(SiriusHipStats, kernel_ms), (int64_t)offsetof(SiriusHipStats, null_drift_median),
        (int64_t)sizeof(SiriusHipImageOpts), (int64_t)sizeof(SiriusHipImageStats),
        (int64_t)offsetof(SiriusHipImageStats, payload_bytes),
    };
    int m = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < m; ++i) out[i] = v[i];
    return m;
}
