/*
 * sirius_oracle.h — CPU FP64 oracle for the geodesic-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the smoke entry and the
 * CPU-baseline leg of bench.py may load it. The product (libsirius_hip.so) never links,
 * loads or calls it, and has no CPU fallback.
 *
 * It restates the reference's CPU path in pure FP64 (the reference rounds several
 * intermediates through float, see PHGD001A.cpp:176,199,549-571 and GTRC001A.cpp:35-38; the
 * oracle does not) and applies the build's trace contract (termination on the Kerr r, exact
 * event location, Page-Thorne emission, exact g-factor). Every function names the reference
 * lines it follows. Pins: tests/test_oracle_*.py.
 */
#ifndef SIRIUS_ORACLE_H
#define SIRIUS_ORACLE_H

#include "../include/sirius_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Metric and its first derivatives at x = (t, x, y, z) (Cartesian; spherical (t,l,th,ph) for
 * Morris-Thorne). g and ginv are 4x4 row-major; dg[lam*16 + mu*4 + nu] = d g_{mu nu}/d x^lam
 * for lam = 0..3 (lam = 0 is always zero: every metric here is stationary). */
void oracle_metric(const SiriusHipMetric* m, const double x[4], double g[16], double ginv[16],
                   double dg[64]);
double oracle_kerr_radius(const SiriusHipMetric* m, double x, double y, double z);
double oracle_horizon(const SiriusHipMetric* m);
double oracle_isco(double a);
double oracle_photon_orbit(double a, int prograde);
/* Page-Thorne zero-torque flux factor f(r) (M = 1), flux F = Mdot f / (4 pi r). */
double oracle_page_thorne_f(double r, double a);
/* T(r) / T_max for a disk on [r_in, r_out] (0 outside or inside the ISCO). */
double oracle_disk_temperature_ratio(double r, double a, double r_in, double r_out);
double oracle_disk_tmax_f_over_r(double a, double r_in, double r_out);
/* Hamiltonian right-hand side: dx = g^{mu nu} p_nu, dp = 1/2 d_mu g_{rs} k^r k^s. Returns p.k */
double oracle_rhs(const SiriusHipMetric* m, const double x[4], const double p[4], double dx[4],
                  double dp[4]);
/* Initial state of pixel (px, py) of a camera: x^mu and p_mu (null, future-directed). */
/* Offsets every later ray's image point by (du, dv) pixels (a test differentiating traces with
 * respect to the image coordinates); 0, 0 restores the pixel centres. */
void oracle_set_subpixel(double du, double dv);
void oracle_camera_state(const SiriusHipCamera* cam, const SiriusHipMetric* m, int px, int py,
                         double x[4], double p[4]);
/* Distance (M) of the pixel's (alpha, beta) from the critical curve, negative inside. */
double oracle_critical_distance(const SiriusHipCamera* cam, const SiriusHipMetric* m, int px,
                                int py);
void oracle_bin_centres(int n_bins, double x_min, double x_max, double* centres);
void oracle_default_opts(SiriusHipTraceOpts* opts, int precision);

/* Traces the rays listed in ray_ids (global ray numbers, camera-major); output slot i belongs
 * to ray_ids[i]. ray_ids == NULL traces every ray and n_ids must equal the total. OpenMP over
 * rays with n_threads threads (0: all cores). Host buffers only. */
int oracle_trace(const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics,
                 int n_metric, const SiriusHipTraceOpts* opts, const int64_t* ray_ids,
                 int64_t n_ids, SiriusHipRayOutput* out, int n_threads, SiriusHipStats* stats);

int oracle_abi_layout(int64_t* out, int n);

/* MLP oracle (FP64): the surrogate network of sirius_hip.h (same parameter layout, inputs,
 * outputs and loss). round_bf16 != 0 rounds every GEMM operand to BF16 (round to nearest
 * even) exactly where the device does; 0 is the plain FP64 network. */
size_t oracle_mlp_param_count(void);
void oracle_mlp_forward(const float* params, const float* x, int n, double* out, int round_bf16);
/* Mean loss over the batch and its gradient w.r.t. params (length param_count). */
double oracle_mlp_loss_grad(const float* params, const float* x, const int32_t* cls,
                            const float* dir, int n, double* grad, int round_bf16);

#ifdef __cplusplus
}
#endif

#endif
