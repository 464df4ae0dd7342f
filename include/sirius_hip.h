/*
 * sirius_hip.h — C-ABI of the MI355X geodesic-tracing backend (libsirius_hip.so).
 *
 * Plain C: opaque handle, plain structs of doubles/ints, plain pointers. No HIP, CUDA or
 * torch types cross this boundary, and no C++ exception escapes it: every entry point that
 * can fail returns an int status (SIRIUS_HIP_OK = 0) and records a message readable with
 * sirius_hip_last_error().
 *
 * The lifecycle/framebuffer half mirrors the OptiX backend's C API one for one
 * (src/Sirius.Render/Acceleration/OptiX/RDOP006A.h:10-39, implemented at
 * src/Sirius.Render/Acceleration/OptiX/RDOP001A.cu:1367-1458); the physics half
 * (sirius_hip_trace_batch and friends) is new: the reference has no batched, per-ray physics
 * entry point.
 *
 * Conventions (G = c = M_unit = 1):
 *  - Metric families follow LaunchConfig::metricFamily
 *    (src/Sirius.Render/Acceleration/Backend/ACIB001A.h:89): 0 Kerr-Schild (M,a,Q,Lambda),
 *    1 Morris-Thorne, 2 Alcubierre.
 *  - Cameras are pinhole cameras as PinholeCamera::generateRay
 *    (src/Sirius.Core/Camera/CMBS001A.h:120-185); the observer sits at Boyer-Lindquist
 *    (r_obs, theta_obs, phi_obs), mapped to Cartesian with the flat map the CPU tracer uses
 *    (src/Sirius.Core/Coordinate/PHCT002A.h:129-140).
 *  - Rays of a batch are numbered camera by camera, row-major inside a camera:
 *    ray = offset(camera) + y * width + x.
 */
#ifndef SIRIUS_HIP_H
#define SIRIUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIRIUS_HIP_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------------------ */
enum {
    SIRIUS_HIP_OK = 0,
    SIRIUS_HIP_E_INVALID = -1,     /* bad argument (null pointer, size, enum)            */
    SIRIUS_HIP_E_DEVICE = -2,      /* HIP runtime error or no usable gfx950 device         */
    SIRIUS_HIP_E_NOMEM = -3,       /* device allocation failed                             */
    SIRIUS_HIP_E_STATE = -4,       /* call out of order (not initialised, no network, ...) */
    SIRIUS_HIP_E_UNSUPPORTED = -5, /* option combination not implemented                   */
    SIRIUS_HIP_E_IO = -6           /* a file could not be written                          */
};

/* ---- enums ---------------------------------------------------------------------------- */
enum { SIRIUS_METRIC_KERR_SCHILD = 0, SIRIUS_METRIC_MORRIS_THORNE = 1, SIRIUS_METRIC_ALCUBIERRE = 2 };
enum { SIRIUS_INTEGRATOR_DP45 = 0, SIRIUS_INTEGRATOR_SYMPLECTIC = 1 };
/* FP32_ESCALATE: FP32 stepping on a compensated state (FP64 initial state kept as high + low
 * FP32 parts, error-free accumulation of every step) with FP64 null-constraint
 * re-normalisation; a ray is re-traced in FP64 from its pixel when it meets an escalation
 * criterion: Kerr-Schild - the options escalate_band, escalate_margin, escalate_disk_edge (all
 * off by default); Morris-Thorne - impact parameter within 10 % of the throat radius;
 * Alcubierre - entering the bubble wall (r_s < R + 3 / sigma) (both always on). */
enum { SIRIUS_PRECISION_FP64 = 0, SIRIUS_PRECISION_FP32_ESCALATE = 1 };
enum { SIRIUS_EMISSION_NONE = 0, SIRIUS_EMISSION_THIN_DISK = 1, SIRIUS_EMISSION_VOLUMETRIC = 2 };
enum { SIRIUS_SCHEDULE_GRID = 0, SIRIUS_SCHEDULE_PERSISTENT = 1 };

/* Per-ray termination class. Values 0-4 keep TraceResult::Outcome's order
 * (src/Sirius.Render/Integration/GTRC001A.h:46-52); 5 is new. */
enum {
    SIRIUS_RAY_ESCAPED = 0,   /* reached r_escape moving outward                           */
    SIRIUS_RAY_CAPTURED = 1,  /* reached r < horizon_factor * r_+                          */
    SIRIUS_RAY_DISK = 2,      /* hit the emitting disk (thin disk)                         */
    SIRIUS_RAY_MAX_STEPS = 3, /* step budget exhausted                                     */
    SIRIUS_RAY_SPIRAL = 4,    /* reserved (reference's SPIRALING); never produced          */
    SIRIUS_RAY_FAILED = 5     /* NaN/Inf state or step size below h_min                    */
};
/* For two-sided wormholes, ESCAPED rays carry the exit side in flags bit 4. */

/* RayResult flags (bit mask) */
enum {
    SIRIUS_FLAG_ESCALATED = 1,   /* the FP32 path handed this ray to FP64                  */
    SIRIUS_FLAG_SURROGATE = 2,   /* class/direction came from the MLP, not the integrator  */
    SIRIUS_FLAG_RENORM = 4,      /* the null constraint was re-imposed at least once       */
    SIRIUS_FLAG_AUDITED = 8,     /* ray belongs to the random audit set                    */
    SIRIUS_FLAG_FAR_SIDE = 16,   /* wormhole: escaped on the far side of the throat        */
    SIRIUS_FLAG_CRITICAL = 32    /* FP32 path: within critical_band of the critical curve  */
};

/* ---- descriptors ---------------------------------------------------------------------- */
typedef struct SiriusHipMetric {
    int32_t family;      /* SIRIUS_METRIC_*                                               */
    int32_t reserved0;
    double M, a, Q, Lambda;             /* Kerr-Schild family, PHMT100A.h:54-67            */
    double throat_b0;                   /* Morris-Thorne throat, b(r) = b0^2 / r           */
    double warp_v, warp_R, warp_sigma;  /* Alcubierre bubble (static, centred at origin)   */
} SiriusHipMetric;

typedef struct SiriusHipCamera {
    double r_obs, theta_obs, phi_obs;   /* observer, Boyer-Lindquist (CMBS001A.h:37-41)    */
    double fov_deg;                     /* vertical field of view (CMBS001A.h:49)          */
    double yaw, pitch, roll;            /* radians (CMBS001A.h:44-46)                      */
    double disk_r_in, disk_r_out;       /* thin disk bounds; disk_r_in <= 0 means ISCO     */
    int32_t width, height;
    int32_t metric_index;               /* index into the metric array                     */
    int32_t reserved0;
} SiriusHipCamera;

typedef struct SiriusHipTraceOpts {
    int32_t integrator;      /* SIRIUS_INTEGRATOR_*                                         */
    int32_t precision;       /* SIRIUS_PRECISION_*                                          */
    int32_t emission;        /* SIRIUS_EMISSION_*                                           */
    int32_t schedule;        /* SIRIUS_SCHEDULE_*                                           */
    double rtol, atol;       /* DP45 tolerances (IntegratorConfig, PHGD001A.h:48-58)        */
    double h_init, h_min, h_max;
    double far_slope;        /* FP32 path: h_max(r) = max(h_max, far_slope * r); 0 = off.
                              * When on, a ray's first step is max(h_init, h_max(r_start))    */
    double symplectic_h;     /* fixed step of the symplectic integrator                     */
    int32_t symplectic_order;/* 2 or 4 (Yoshida composition)                                */
    int32_t max_steps;       /* step attempts per ray                                       */
    double r_escape;         /* escape radius                                               */
    double horizon_factor;   /* capture at r < horizon_factor * r_+                         */
    int32_t n_bins;          /* spectral bins, log-spaced in h nu / k T_max                 */
    int32_t use_surrogate;   /* 1: MLP far-field surrogate with integrator fallback         */
    double x_min, x_max;     /* bin range in h nu / k T_max                                 */
    double escalate_margin;  /* FP32 path: escalate at a radial turning point inside
                              * r < escalate_margin * r_ph(retro); 0 = off (default)          */
    double escalate_band;    /* FP32 path: rays within this distance (M) of the critical
                              * curve are traced in FP64 from the start on a second stream
                              * (Kerr-Schild only); 0 = off (default)                          */
    double surrogate_band;   /* fallback band around the critical curve (M)                 */
    double surrogate_pmin;   /* fallback when max class probability is below this           */
    double audit_fraction;   /* fraction of surrogate rays re-traced in FP64                */
    uint64_t audit_seed;
    int32_t block_size;      /* 0: autotuned                                                */
    int32_t blocks_per_cu;   /* 0: autotuned                                                */
    int32_t refill_min;      /* persistent queue: refill a wave when >= this many lanes idle*/
    int32_t reserved0;
    /* volumetric emission (emitter frame): j_nu = rho(x) x^3 / (e^{x / tau_v(r)} - 1),
     * rho = j0 r^-2 exp(-z^2 / (2 (h_over_r r)^2)) W(r), tau_v = (6 / r)^{3/4},
     * alpha = kappa rho; W is a C2 (quintic) smoothstep window, 0 at r_lo and vol_r_max and 1
     * on [1.1 r_lo, 0.9 vol_r_max], r_lo = max(vol_r_min, the camera's disk r_in, by default the
     * ISCO); emitters are on Keplerian (Kerr-Newman) circles */
    double vol_j0, vol_h_over_r, vol_kappa, vol_r_min, vol_r_max, vol_tau_max;
    double escalate_disk_edge; /* FP32 path, thin disk: a disk crossing whose radius lies within
                              * this fraction of r_in (|r/r_in - 1| < edge) is re-traced in FP64,
                              * where the emission's steep edge magnifies position error; 0 =
                              * off (default: the compensated FP32 state meets the intensity
                              * tolerance there)                                             */
    double escalate_rtol;    /* relative tolerance of escalated FP64 re-traces (atol scales with
                              * it); they keep the FP32 path's step caps. 0 (default): the
                              * path's rtol, except 3e-10 for the warp bubble's wall crossers
                              * and 1e-8 for the wormhole's near-throat rays, which the FP32
                              * tolerance does not control                                     */
    double critical_band;    /* FP32 path, Kerr-Schild: rays within this distance (M) of the
                              * critical curve are flagged SIRIUS_FLAG_CRITICAL; the accuracy
                              * contract exempts them and the inner-edge escalation skips them */
    double step_radial_exp;  /* DP45 step proposal after an accepted step that moved inward
                              * (r1 < r0): h_next = h fac(err) (r1 / r0)^exp, anticipating the
                              * local error's growth toward the hole. 0 = off (the reference's
                              * controller, the FP64 default); FP32 default 2              */
    double surrogate_err_max;/* surrogate: answer a ray only where the escape-direction error the
                              * calibration measured on held-out rays of its bin is at most this
                              * (rad); default 5e-5                                            */
    double surrogate_margin; /* surrogate: the analytic candidate test refuses a ray whose
                              * reference orbit crosses the equatorial plane within this distance
                              * (M) of the thin disk's [r_in, r_out]; default 2.5              */
    double precision_audit;  /* FP32 path: this fraction of the traced rays outside the critical
                              * band (not flagged SIRIUS_FLAG_CRITICAL: the rays the accuracy
                              * contract covers; chosen by a hash of audit_seed and the ray index)
                              * is re-traced after the batch by the plain FP64 integrator at the
                              * FP64 defaults' tolerances and compared: the stats' precision_*
                              * fields estimate the call's error against FP64. 0 = off (default) */
} SiriusHipTraceOpts;

/* Structure-of-arrays result buffers. Every array holds n_rays entries, except intensity
 * (n_bins * n_rays, bin-major: intensity[b * n_rays + ray]) and state (8 * n_rays, optional).
 * on_device != 0: non-null pointers are device memory written in place, and null ones leave
 * that result in library-owned device memory (fetch it with sirius_hip_fetch_results);
 * on_device == 0: non-null pointers are host memory (the library stages through device
 * scratch and copies back before returning), null ones are skipped. */
typedef struct SiriusHipRayOutput {
    uint8_t* cls;        /* SIRIUS_RAY_*                                                   */
    uint8_t* flags;      /* SIRIUS_FLAG_*                                                  */
    float* theta;        /* escape direction polar angle of k^i (rad)                       */
    float* phi;          /* escape direction azimuth of k^i (rad)                           */
    float* lambda;       /* affine length travelled                                        */
    float* disk_r;       /* Boyer-Lindquist radius of the disk hit                          */
    float* redshift;     /* g = nu_obs / nu_emit at the disk hit                            */
    float* null_drift;   /* |g_{mu nu} k^mu k^nu| at termination (evaluated in FP64)        */
    uint32_t* steps;     /* step attempts (accepted + rejected)                             */
    float* intensity;    /* n_bins * n_rays                                                 */
    double* state;       /* optional final (x^mu, p_mu), 8 * n_rays, ray-major              */
    int32_t on_device;
    int32_t reserved0;
} SiriusHipRayOutput;

typedef struct SiriusHipStats {
    uint64_t rays;
    uint64_t steps_accepted;
    uint64_t steps_rejected;
    uint64_t steps_landing;      /* event-location steps (disk, horizon, escape sphere)     */
    uint64_t rays_escalated;     /* FP32 rays finished in FP64                              */
    uint64_t rays_surrogate;     /* rays answered by the MLP                                */
    uint64_t rays_fallback;      /* rays the gate sent back to the integrator               */
    uint64_t rays_audited;
    uint64_t steps_bulk;         /* step attempts made by the bulk trace kernel             */
    uint64_t steps_escalated;    /* step attempts made by the FP64 escalation kernel        */
    double kernel_ms;            /* HIP-event time of the bulk trace kernel                 */
    double escalate_ms;          /* HIP-event time of the FP64 escalation kernel            */
    double surrogate_ms;         /* HIP-event time of the MLP inference + gate              */
    double total_ms;             /* HIP-event time of the whole batch                       */
    double flops;                /* algorithmic flops of the bulk kernel (fixed per step)   */
    double flops_escalated;      /* algorithmic flops of the escalation kernel              */
    double null_drift_max;       /* largest |g^{mu nu} p_mu p_nu| of the final states         */
    double null_drift_median;    /* its median over the batch, to 1/16 decade (log10 bins;
                                  * 0 when the middle ray's drift is below 1e-20)             */
    double audit_ms;             /* HIP-event time of the surrogate audit (FP64 re-traces)   */
    double audit_dir_err_max;    /* largest escape-direction error of audited answers (rad) */
    uint64_t audit_class_mismatch; /* audited answers whose class the integrator disputes    */
    double surrogate_flops;      /* algorithmic flops of the network inference               */
    uint64_t rays_candidates;    /* rays the surrogate's analytic test passed to the network
                                    (only rays in a calibration bin that answers)            */
    uint64_t rays_gated;         /* candidates the calibrated error gate left to the integrator */
    uint64_t precision_audited;  /* rays of the FP64 precision audit (precision_audit)        */
    uint64_t precision_class_mismatch; /* audited rays whose FP32 class differs from FP64's    */
    double precision_dir_err_max;      /* escape-direction error of the audited rays (rad):   */
    double precision_dir_err_median;   /* largest and median, both classes escaped            */
    double precision_int_err_max;      /* relative intensity error of audited disk rays (the  */
    double precision_int_err_median;   /* largest bin error over the pixel's brightest bin)   */
    double precision_audit_ms;   /* HIP-event time of the precision audit                    */
} SiriusHipStats;

typedef void* SiriusHipHandle;

/* ---- lifecycle: mirrors sirius_optix_* (RDOP006A.h:14-19) ----------------------------- */
SiriusHipHandle sirius_hip_create(void);
void sirius_hip_destroy(SiriusHipHandle h);
/* Allocates a width*height float4 framebuffer (IAccelerator::initialise, ACIB001A.h:160). */
int sirius_hip_initialize(SiriusHipHandle h, int width, int height);
void sirius_hip_cleanup(SiriusHipHandle h);
int sirius_hip_is_initialized(SiriusHipHandle h);
const char* sirius_hip_last_error(SiriusHipHandle h);
int sirius_hip_abi_version(void);
/* Device name, memory, CU count, FP64 support (DeviceCapabilities, ACIB001A.h:50-59). */
int sirius_hip_device_info(SiriusHipHandle h, char* name, int name_len, size_t* total_mem,
                           size_t* free_mem, int* cu_count, int* max_threads_per_block);

/* ---- rendering: mirrors sirius_optix_launch / get_frame_buffer (RDOP006A.h:22-29) ------ */
/* Framebuffer shading of sirius_hip_launch: linear RGB radiance, as RenderSession fills its display
 * buffer before post-processing (SNRS001A.cpp:480-560 per pixel, shadeDiskHit :315-434):
 *  - each pixel averages a g x g grid of rays, g = floor(sqrt(samples_per_pixel)), through the
 *    points ((sx + 1/2) / g, (sy + 1/2) / g) of the pixel;
 *  - thin-disk hit (emission SIRIUS_EMISSION_THIN_DISK): emitted temperature T_e(r) =
 *    disk_temperature * (F(r) / F_max)^(1/4) with the Page-Thorne flux F (temperature_model 1,
 *    Novikov-Thorne) or disk_temperature * (r_in / r)^(3/4) (0, Shakura-Sunyaev); with the ray's g
 *    and t = T_e g / 30000 K, colour = blackbody RGB of clamp(T_e g, 1000 K, 1e5 K) (brightest channel
 *    1; PHSP001A.h:181-205) * t^4 g^4 (ColorModes TrueColor, PHSC001A.h:352-366) * the channel's limb
 *    darkening (1 + u cos i) / (1 + u) (PHSP001A.h:306-322; cos i = |k_z| / |k| at the hit) * g^4 *
 *    disk_boost;
 *  - escaped: the background image (or black), plus star light when a catalogue is installed;
 *  - captured: black; ended on max_steps or failed: grey max_steps_level.
 * The reference's photon-ring magnification and boosts (heuristics of its CPU tracer,
 * GTRC001A.cpp:395-431, 528-546) are not applied. */
typedef struct SiriusHipShading {
    int32_t samples_per_pixel;  /* LaunchConfig::samplesPerPixel (default 1)                     */
    int32_t temperature_model;  /* 0 Shakura-Sunyaev, 1 Novikov-Thorne (LaunchConfig default 1)  */
    double disk_temperature;    /* K (LaunchConfig::diskTemperature, default 50000)              */
    double disk_boost;          /* DISK_INTENSITY_BOOST (SNRS001A.cpp:35), default 5              */
    double max_steps_level;     /* MAX_STEPS_BRIGHTNESS (SNRS001A.cpp:41), default 0.01           */
} SiriusHipShading;
void sirius_hip_default_shading(SiriusHipShading* s);
/* Shading of later sirius_hip_launch calls (the defaults until called). */
int sirius_hip_set_shading(SiriusHipHandle h, const SiriusHipShading* s);
/* Traces one camera and shades the RGBA framebuffer (synchronous, like RDOP001A.cu:1044). With
 * g x g samples per pixel the held results (sirius_hip_fetch_results) are those of the camera of
 * g W x g H pixels whose rays the frame averages. */
int sirius_hip_launch(SiriusHipHandle h, const SiriusHipCamera* cam, const SiriusHipMetric* metric,
                      const SiriusHipTraceOpts* opts);
int sirius_hip_resize(SiriusHipHandle h, int width, int height);
/* Host RGBA float pointer (width*height*4); the caller does not own it (ACIB001A.h:169-170). */
float* sirius_hip_get_frame_buffer(SiriusHipHandle h);
size_t sirius_hip_get_frame_buffer_size(SiriusHipHandle h);
/* Equirectangular RGBA8 background (IAccelerator::uploadBackground, ACIB001A.h:176). */
int sirius_hip_upload_background(SiriusHipHandle h, const unsigned char* rgba, int width, int height);
int sirius_hip_synchronize(SiriusHipHandle h);
/* Copies the results of the most recent trace_batch or launch held in library device memory
 * into host buffers (host pointers; null ones skipped). SIRIUS_HIP_E_STATE once a later call
 * has replaced the library's result arrays. Only library-owned arrays are held: a result that
 * trace_batch wrote into the caller's device array (on_device) is not copied here (its host
 * array is left untouched), and sirius_hip_write_rays leaves it out of the file. */
int sirius_hip_fetch_results(SiriusHipHandle h, SiriusHipRayOutput* host_out);
/* Size of those results: rays and spectral bins (what fetch_results' host arrays must hold).
 * SIRIUS_HIP_E_STATE when there are none. Has no counterpart in the reference. */
int sirius_hip_last_result_size(SiriusHipHandle h, int64_t* n_rays, int32_t* n_bins);

/* ---- physics: batched trace (new; no reference counterpart) --------------------------- */
/* Fills opts with the defaults of the given precision (FP64 reference contract or FP32). */
void sirius_hip_default_opts(SiriusHipTraceOpts* opts, int precision);
/* Traces every pixel of n_cam cameras. metrics[cam.metric_index] gives each camera's
 * spacetime. Returns SIRIUS_HIP_OK or an error code; stats may be null. */
int sirius_hip_trace_batch(SiriusHipHandle h, const SiriusHipCamera* cams, int n_cam,
                           const SiriusHipMetric* metrics, int n_metric,
                           const SiriusHipTraceOpts* opts, SiriusHipRayOutput* out,
                           SiriusHipStats* stats);
/* sirius_hip_trace_batch without allocation: the same kernels, with every device scratch
 * buffer carved from one fixed array of 256 MiB that the library holds per GPU until the process
 * ends, so the call makes no hipMalloc / hipFree (nor any other device or pinned-host allocation).
 * Has no counterpart in the reference; it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 *  - out->on_device must be 1: results go to the caller's device arrays. A null array is
 *    computed in the scratch and discarded; sirius_hip_fetch_results then has nothing to fetch.
 *  - Limit: the batch's scratch (88 bytes per ray with every result array supplied except state;
 *    24 with state supplied too; more for volumetric emission and for arrays left null) must fit
 *    in 256 MiB, else SIRIUS_HIP_E_INVALID before any work. A 1920x1080 frame fits.
 *  - use_surrogate is not supported (SIRIUS_HIP_E_UNSUPPORTED).
 *  - Every call on one GPU, from any handle, shares the array and leaves its contents behind
 *    (each call initialises what it reads); calls that overlap on one GPU race. */
int sirius_hip_trace_batch_noalloc(SiriusHipHandle h, const SiriusHipCamera* cams, int n_cam,
                                   const SiriusHipMetric* metrics, int n_metric,
                                   const SiriusHipTraceOpts* opts, SiriusHipRayOutput* out,
                                   SiriusHipStats* stats);
/* Distance (M) of each ray from the critical curve in the (alpha, beta) plane, measured
 * radially from the curve's centre (alpha_0, 0): rho - rho_c(psi), negative inside the shadow.
 * alpha = -L/(E sin th), beta = p_th/E at the ray's start. Rays are all pixels of the cameras,
 * output n_rays floats (device or host). */
int sirius_hip_critical_distance(SiriusHipHandle h, const SiriusHipCamera* cams, int n_cam,
                                 const SiriusHipMetric* metrics, int n_metric, float* out,
                                 int on_device);

/* ---- autotuner (new) ------------------------------------------------------------------- */
/* Times the persistent kernel over a sweep of (block_size, blocks_per_cu, refill_min) on a
 * sample of the batch and writes the fastest choice into opts. */
int sirius_hip_autotune(SiriusHipHandle h, const SiriusHipCamera* cams, int n_cam,
                        const SiriusHipMetric* metrics, int n_metric, SiriusHipTraceOpts* opts,
                        double* best_ms);

/* ---- MLP far-field surrogate (new) ------------------------------------------------------ */
/* Network: 5 inputs -> 4 x 256 GELU (tanh form) -> 6 outputs: 3 class logits (escaped,
 * captured, disk) and the escape direction's residual over an analytic reference,
 * (d_esc - d_ref) / SIRIUS_MLP_DIR_SCALE. d_ref is the exact exit direction (at r_escape) of the
 * Schwarzschild ray with the same initial coordinate direction starting from the initial
 * position displaced by 1.1 a e_phi (e_phi: the observer's azimuthal unit vector; to first order
 * in a, frame dragging acts as such a displacement of the lens): at a = Q = Lambda = 0 the
 * integrator's own answer, so the network learns what spin beyond that, charge and Lambda add. Inputs: a, theta_obs / (pi/2), (r_obs - 60) / 40, alpha / 15, beta / 15 (alpha,
 * beta: the ray's impact parameters on the observer's sky, M).
 * The network is asked only about candidates: rays outside surrogate_band of the critical curve
 * and outside the shadow whose reference orbit scatters and crosses the equatorial plane only
 * outside the thin disk widened by surrogate_margin (the reference's own test, analytic), inside
 * the training cameras' field of view (image-plane box), and in a calibration bin (spin x
 * critical distance x observer radius) whose held-out rays showed no misclassification and a
 * largest direction error <= surrogate_err_max, at a critical distance no larger than those
 * held-out rays reached. A candidate is answered when its most probable class is escaped with
 * probability >= surrogate_pmin; every other ray is integrated.
 * Parameters are FP32 masters on the device, laid out layer by layer as W_l (out x in,
 * row-major) then b_l. The hidden layers run on BF16 MFMA with FP32 accumulation (operands
 * rounded to BF16: activations, weights and pre-activation gradients); the output head (256 ->
 * 6) is FP32. Loss per sample: softmax cross-entropy + SIRIUS_MLP_DIR_WEIGHT * |residual
 * error|^2 for escaped rays; the batch loss is the mean. */
#define SIRIUS_MLP_IN 5
#define SIRIUS_MLP_HIDDEN 256
#define SIRIUS_MLP_LAYERS 4
#define SIRIUS_MLP_OUT 6
#define SIRIUS_MLP_DIR_WEIGHT 100.0
#define SIRIUS_MLP_DIR_SCALE 0.005
size_t sirius_hip_mlp_param_count(void);
/* Deterministic initialisation from a seed (uniform +-sqrt(6 / fan_in) weights, zero biases). */
int sirius_hip_mlp_init(SiriusHipHandle h, uint64_t seed);
int sirius_hip_mlp_get_params(SiriusHipHandle h, float* host_params, size_t n);
int sirius_hip_mlp_set_params(SiriusHipHandle h, const float* host_params, size_t n);
/* Forward pass on n samples x[n*5] (device), outputs y[n*6] (device). */
int sirius_hip_mlp_forward(SiriusHipHandle h, const float* x, float* y, int n);
/* Loss and parameter gradient for a batch; targets cls[n] (int32: 0 escaped, 1 captured,
 * 2 disk) and dir[n*3] (residuals, read for escaped rays); grad is FP32, length param_count
 * (device). The mean loss is returned in *loss (host). */
int sirius_hip_mlp_loss_grad(SiriusHipHandle h, const float* x, const int32_t* cls,
                             const float* dir, int n, float* grad, float* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) over a batch on device buffers; *loss (host)
 * receives the batch loss before the update. */
int sirius_hip_mlp_train_step(SiriusHipHandle h, const float* x, const int32_t* cls,
                              const float* dir, int n, float lr, float* loss);
/* Network inputs x[n*5] and reference exit directions ref[n*3] of every ray of the cameras
 * (device buffers; Kerr-Schild cameras; ref is zero for rays the candidate test refuses, with
 * the FP32 defaults' surrogate_band, surrogate_margin and r_escape). */
int sirius_hip_surrogate_features(SiriusHipHandle h, const SiriusHipCamera* cams, int n_cam,
                                  const SiriusHipMetric* metrics, int n_metric, float* x, float* ref);
/* Trains the network on the candidate rays of the cameras, labelled by the integrator with
 * label_opts (FP64; its surrogate_band, surrogate_margin and r_escape define the candidates):
 * inputs and targets are built on the device; a seeded 1 in 4 of the candidates is held out,
 * the rest trained on for `epochs` passes of Adam over batches of `batch` rays in a seeded
 * random order, at learning rate lr with a cosine decay to lr/20; then the held-out rays
 * calibrate the error gate (label_opts' surrogate_pmin). *loss (host) receives the mean batch
 * loss of the last epoch. Uses the current parameters (sirius_hip_mlp_init first for a fresh
 * network). epochs = 0 trains nothing and only calibrates the current parameters, receiving
 * *loss = 0: the calibration of parameters restored by sirius_hip_mlp_set_params (which clears
 * it), whose training cameras and seed give the same held-out rays again. */
int sirius_hip_surrogate_train(SiriusHipHandle h, const SiriusHipCamera* cams, int n_cam,
                               const SiriusHipMetric* metrics, int n_metric,
                               const SiriusHipTraceOpts* label_opts, int epochs, int batch,
                               float lr, uint64_t seed, float* loss);
/* The error gate's calibration, per bin (*n_bins of them; capacity must be at least that):
 * held-out rays the network would answer, how many of them it misclassified, and the largest
 * escape-direction error among the rest (rad). All zero before sirius_hip_surrogate_train, and
 * after new parameters (an uncalibrated network answers nothing). */
int sirius_hip_surrogate_calibration(SiriusHipHandle h, float* max_err, uint32_t* answered,
                                     uint32_t* misclassified, int capacity, int* n_bins);

/* ---- procedural star field (celestial spheres) ---------------------------------------- */
/* One catalogue star, field for field the reference's StarEntry
 * (src/Sirius.Core/Environment/PHSF001A.h:38-46): unit direction, distance (pc), apparent
 * magnitude, B-V colour index, effective temperature (K). */
typedef struct SiriusHipStar {
    float dir[3];
    float distance_pc;
    float magnitude;
    float color_bv;
    float temperature_K;
    float reserved;
} SiriusHipStar;

/* The generator and sampler parameters of the reference's StarfieldConfig (PHSF001A.h:85-121);
 * its galactic-model fields, which its generator does not read, are left out. */
typedef struct SiriusHipStarfield {
    uint32_t star_count;      /* candidates drawn; those brighter than magnitude_limit are kept */
    uint32_t seed;            /* std::mt19937 seed                                              */
    float min_distance_pc, max_distance_pc;
    float magnitude_limit;
    float aperture_mm;        /* depth-of-field blur (0: none)                                  */
    float focus_distance_pc;
    float brightness_scale;
    int32_t enable_dof;
    int32_t reserved;
} SiriusHipStarfield;

/* The reference's defaults (StarfieldConfig, PHSF001A.h:85-98). */
void sirius_hip_starfield_default(SiriusHipStarfield* cfg);
/* StarfieldGenerator::generate (PHSF001A.h:133-151), on the host: the catalogue for cfg (after
 * the reference's validate() clamps); writes at most `capacity` stars, *count receives the
 * catalogue's size. No device needed. */
int sirius_hip_starfield_generate(const SiriusHipStarfield* cfg, SiriusHipStar* stars, int capacity, int* count);
/* Installs a catalogue as celestial sphere `side` (0: the sky, or the near side of a wormhole;
 * 1: a wormhole's far side). sirius_hip_launch then adds each escaped ray's star light
 * (StarfieldGenerator::sampleStarfield, PHSF001A.h:153-201, observer at the origin) to the
 * framebuffer. n = 0 removes the sphere. Has no counterpart in the reference's C API, whose
 * star field lives in the CPU session. */
int sirius_hip_set_starfield(SiriusHipHandle h, int side, const SiriusHipStarfield* cfg, const SiriusHipStar* stars,
                             int n);
/* Star light of sphere `side` at n unit directions dirs[3n] (host arrays), rgb[3n] out, computed
 * on the device by the kernel the shading uses. */
int sirius_hip_sample_starfield(SiriusHipHandle h, int side, const float* dirs, int n, float* rgb);

/* ---- Ray bundles ---------------------------------------------------------------------------
 * The geodesic deviation of each pixel's ray: two deviation vectors, the derivatives of the ray's
 * state with respect to the image coordinates (u, v), carried along the ray as the linearised
 * flow (the geodesic deviation equation in Hamiltonian form), and from them the derivative of
 * the ray's outcome: the Jacobian of the escape direction (theta, phi) for an escaped ray, in rad
 * per pixel, or of the equatorial hit point (X, Y) for a ray ending on the thin disk, in M per
 * pixel; and the beam, the image of a circle of radius size pixels at the camera: semi-axes and
 * the major axis's angle (on the sky from e_theta towards e_phi, in radians; on the disk from
 * the X axis). Replaces the ray bundle of the reference's device kernel
 * (RDOP002A.cu:4239-4354, 5014-5047, 5364-5439), whose parameters are LaunchParams::rayBundle
 * (RDOP003A.h:1063-1069). */
#define SIRIUS_BUNDLE_TERMINATED 1 /* beam unusable (captured, too elongated or too small): point-sampled */

typedef struct SiriusHipRayBundle {
    int32_t enabled;           /* sirius_hip_launch averages star light over each pixel's beam      */
    int32_t use_full_riemann;  /* accepted for the reference's field; the deviation is always the
                                  exact linearised flow                                             */
    float initial_size;        /* beam radius at the camera, pixels; default 0.2887 (the standard
                                  deviation of a unit box pixel)                                    */
    float min_bundle_size;     /* beams whose semi-major axis is below this are terminated
                                  (default 1e-6, the reference's)                                   */
    float max_ellipse_ratio;   /* beams more elongated than this are terminated (default 100)       */
    int32_t reserved0;
} SiriusHipRayBundle;

typedef struct SiriusHipBundleOutput {
    uint8_t* cls;        /* SIRIUS_RAY_* of the bundle's central ray                          */
    uint8_t* flags;      /* SIRIUS_BUNDLE_TERMINATED                                          */
    float* jac;          /* 4 * n_rays, component-major: da/du, da/dv, db/du, db/dv           */
    float* beam;         /* 3 * n_rays, component-major: semi-major, semi-minor, orientation  */
    int32_t on_device;   /* as SiriusHipRayOutput.on_device                                   */
    int32_t reserved0;
} SiriusHipBundleOutput;

void sirius_hip_default_ray_bundle(SiriusHipRayBundle* b);
/* Bundle parameters of later sirius_hip_launch calls (enabled = 0: point-sampled stars). */
int sirius_hip_set_ray_bundle(SiriusHipHandle h, const SiriusHipRayBundle* b);
/* Traces every pixel's ray bundle in FP64 (DP45 with opts' tolerances, step limits and
 * far-field cap; precision is ignored) and writes the bundle outputs; beams use the parameters
 * of sirius_hip_set_ray_bundle (its defaults if never set). */
int sirius_hip_trace_bundle(SiriusHipHandle h, const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics,
                            int n_metric, const SiriusHipTraceOpts* opts, SiriusHipBundleOutput* out);

/* ---- Image writers ----------------------------------------------------------------------
 * A float RGBA frame (the framebuffer, or a caller's array) is converted on the device and
 * packed into a file on the host. Each format decodes to exactly the pixels of the reference
 * writer named beside it, for the same float RGBA input. */
#define SIRIUS_IMAGE_PPM 0      /* BufferWriter::writePPM (BFIO001A.h:44-68): ACES tonemap of
                                   value x exposure, 1/2.2 gamma, truncated to 8 bits; P6 RGB  */
#define SIRIUS_IMAGE_PNG 1      /* PNGWriter::writeRGBA (OUPN001A.cpp:51-72): sRGB 8-bit RGBA */
#define SIRIUS_IMAGE_EXR 2      /* EXRWriter::writeEXR, RGBA (OUEW001A.cpp:109-178,181-190):
                                   channels A B G R as half (tinyexr's rounding), ZIP blocks of
                                   16 scanlines                                               */
#define SIRIUS_IMAGE_PFM 3      /* EXRWriter::writePFM (OUEW001A.h:110-126): RGB float, rows
                                   bottom to top, little-endian                               */
#define SIRIUS_IMAGE_RAW 4      /* BufferWriter::writeRaw (BFIO001A.h:71-86): int32 width,
                                   height, 4, then the RGBA floats                             */
#define SIRIUS_IMAGE_PPM_SRGB 5 /* EXRWriter::writePPM (OUEW001A.h:97-107): sRGB 8-bit, P6 RGB */

typedef struct SiriusHipImageOpts {
    int32_t format;      /* SIRIUS_IMAGE_*                                                    */
    float exposure;      /* SIRIUS_IMAGE_PPM only (writePPM's exposure); default 1            */
    int32_t zlib_level;  /* PNG and EXR: 0..9, or -1 for zlib's default (6)                   */
    int32_t threads;     /* host threads that compress; 0: min(16, hardware threads)          */
} SiriusHipImageOpts;

typedef struct SiriusHipImageStats {
    double convert_ms;       /* HIP-event time of the device conversion kernels               */
    double copy_ms;          /* device -> host copy of the converted payload                 */
    double pack_ms;          /* host compression and container                                */
    double total_ms;         /* the whole call                                                */
    uint64_t payload_bytes;  /* converted bytes copied from the device                       */
    uint64_t file_bytes;
} SiriusHipImageStats;

void sirius_hip_default_image_opts(SiriusHipImageOpts* o, int format);
/* Encodes a frame into library-owned memory: *data (valid until the next encode on this handle
 * or its destruction) and *size. rgba null: the framebuffer of sirius_hip_launch (width and
 * height 0 or its size); otherwise width x height RGBA floats, device memory if on_device,
 * else host memory. stats may be null. Has no counterpart in the reference, whose writers take
 * a host array and a path (below). */
int sirius_hip_encode_image(SiriusHipHandle h, const float* rgba, int width, int height, int on_device,
                            const SiriusHipImageOpts* o, const unsigned char** data, size_t* size,
                            SiriusHipImageStats* stats);
/* The same, written to `path` (PNGWriter::writeRGBA, EXRWriter::writeEXR, BufferWriter::writePPM
 * and their kin; like them it fails for an empty frame or null pixels). */
int sirius_hip_write_image(SiriusHipHandle h, const char* path, const float* rgba, int width, int height,
                           int on_device, const SiriusHipImageOpts* o, SiriusHipImageStats* stats);

/* ---- Ray result dump ------------------------------------------------------------------------
 * The results of the last trace_batch or launch, with the cameras, metrics and options that made
 * them, in one file, for audits that re-trace sampled rays on the CPU. Has no counterpart in
 * the reference. Layout (little-endian):
 *   0   char[8]  "SIRRAYS1"
 *   8   uint32   header bytes (128)          12  uint32  n_bins
 *   16  uint64   n_rays                      24  uint32  fields (SIRIUS_RAYS_* bits)
 *   28  int32    n_cameras                   32  uint32  sizeof(SiriusHipCamera)
 *   36  int32    n_metrics                   40  uint32  sizeof(SiriusHipMetric)
 *   44  uint32   sizeof(SiriusHipTraceOpts), 0 without options; the rest of the header is zero
 * then the camera, metric and option records, and then each field present, in bit order, each
 * starting on an 8-byte boundary: n_rays entries, intensity n_bins x n_rays (bin-major), state
 * n_rays x 8 doubles. */
#define SIRIUS_RAYS_CLS (1u << 0)        /* uint8  */
#define SIRIUS_RAYS_FLAGS (1u << 1)      /* uint8  */
#define SIRIUS_RAYS_THETA (1u << 2)      /* float  */
#define SIRIUS_RAYS_PHI (1u << 3)        /* float  */
#define SIRIUS_RAYS_LAMBDA (1u << 4)     /* float  */
#define SIRIUS_RAYS_DISK_R (1u << 5)     /* float  */
#define SIRIUS_RAYS_REDSHIFT (1u << 6)   /* float  */
#define SIRIUS_RAYS_NULL_DRIFT (1u << 7) /* float  */
#define SIRIUS_RAYS_STEPS (1u << 8)      /* uint32 */
#define SIRIUS_RAYS_INTENSITY (1u << 9)  /* float  */
#define SIRIUS_RAYS_STATE (1u << 10)     /* double */
#define SIRIUS_RAYS_ALL 0x7ffu
int sirius_hip_write_rays(SiriusHipHandle h, const char* path, uint32_t fields, const SiriusHipCamera* cams,
                          int n_cams, const SiriusHipMetric* metrics, int n_metrics, const SiriusHipTraceOpts* opts);

#ifdef __cplusplus
}
#endif

#endif /* SIRIUS_HIP_H */
