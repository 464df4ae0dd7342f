// sirius_hip_api.cpp — the extern "C" boundary of libsirius_hip.so (include/sirius_hip.h).
//
// Lifecycle and framebuffer calls mirror the OptiX backend's C API
// (src/Sirius.Render/Acceleration/OptiX/RDOP006A.h:10-39; implementation
// src/Sirius.Render/Acceleration/OptiX/RDOP001A.cu:1367-1458): opaque handle, bool-like
// status, no exception crosses the boundary. The physics entry point sirius_hip_trace_batch
// is new. One device, one stream per handle (RDOP001A.cu:621,631).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sirius_hip.h"
#include "geodesic.hpp"
#include "mlp.hpp"
#include "sirius_internal.hpp"

namespace sirius {

// ---- kernels defined in trace_kernels.hip / util_kernels.hip ----
hipError_t launch_fill(uint32_t* p, int n, uint32_t v, hipStream_t st);
hipError_t launch_finish(int family, int emission, const CamConst* cams, const TraceParams& tp, const DevOut& o,
                         hipStream_t st);
hipError_t launch_trace_persistent(bool fp64, int family, int emission, bool multi, const CamConst* cams,
                                   const TraceParams& tp,
                                   const DevOut& o, const int32_t* list, const uint32_t* list_count, int64_t n,
                                   uint32_t* head, uint32_t* esc_count, int32_t* esc_list,
                                   unsigned long long* attempts_out, int grid, int block, int refill_min,
                                   uint32_t flag_or, hipStream_t st);
hipError_t launch_symplectic(bool fp64, int family, const CamConst* cams, const TraceParams& tp, const DevOut& o,
                             int block, hipStream_t st);
int persistent_blocks_per_cu(bool fp64, int family, int emission, bool multi, int block);
hipError_t launch_reduce_stats(const DevOut& o, int64_t n, unsigned long long* acc, uint32_t* hist,
                               hipStream_t st);
hipError_t launch_critical_distance(const CamConst* cams, int n_cam, const float* tabs, int64_t n, float* out,
                                    hipStream_t st);
hipError_t launch_gate(const CamConst* cams, int n_cam, const float* tabs, int64_t n, double band, double crit_band,
                       double long_band, int32_t* list_a, int32_t* list_b, int32_t* list_l, uint32_t* counts,
                       uint8_t* flags, hipStream_t st);
hipError_t launch_sur_audit(const CamConst* cams, const TraceParams& tp, const DevOut& oa, const DevOut& o,
                            const int32_t* list, const uint32_t* count, int64_t cap, unsigned int* acc,
                            hipStream_t st);
hipError_t launch_precision_select(const DevOut& o, int64_t n, double fraction, uint64_t seed, int32_t* list,
                                   uint32_t* count, hipStream_t st);
hipError_t launch_precision_compare(const CamConst* cams, const TraceParams& tp, const DevOut& oa, const DevOut& o,
                                    const int32_t* list, const uint32_t* count, int64_t cap, unsigned int* acc,
                                    unsigned int* hist, hipStream_t st);

// Trace parameters of the plain FP64 integrator at the FP64 defaults' tolerances and step limits
// (the contract the audits compare against), for the batch of tp.
static TraceParams fp64_reference_params(const TraceParams& tp) {
    SiriusHipTraceOpts d;
    sirius_hip_default_opts(&d, SIRIUS_PRECISION_FP64);
    TraceParams t = tp;
    t.rtol = d.rtol;
    t.atol = d.atol;
    t.h_init = d.h_init;
    t.h_min = d.h_min;
    t.h_max = d.h_max;
    t.far_slope = d.far_slope;
    t.step_rexp = d.step_radial_exp;
    t.max_steps = d.max_steps;
    t.escalate = 0;
    t.gate_flags = 0;
    t.spectrum32 = 0;
    t.first_list = nullptr;
    t.first_count = nullptr;
    return t;
}

// ---------------------------------------------------------------------------------------
// Host-side closed forms (FP64), shared with the camera constant setup.
// ---------------------------------------------------------------------------------------
double isco(double a_star) {
    double a = std::fabs(a_star);
    if (a > 1) a = 1;
    double z1 = 1 + std::cbrt(1 - a * a) * (std::cbrt(1 + a) + std::cbrt(1 - a));
    double z2 = std::sqrt(3 * a * a + z1 * z1);
    double s = std::sqrt((3 - z1) * (3 + z1 + 2 * z2));
    return a_star >= 0 ? 3 + z2 - s : 3 + z2 + s;
}

double photon_orbit(double a, bool prograde) {
    double arg = std::clamp(prograde ? -a : a, -1.0, 1.0);
    return 2.0 * (1.0 + std::cos((2.0 / 3.0) * std::acos(arg)));
}

static double pt_f_host(const CamConst& c, double r) {
    double x = std::sqrt(r), x0 = c.pt_x0;
    if (x <= x0) return 0.0;
    double br = x - x0 - 1.5 * c.a * std::log(x / x0);
    for (int i = 0; i < 3; ++i) br -= c.pt_c[i] * std::log((x - c.pt_x[i]) / (x0 - c.pt_x[i]));
    return 1.5 / (x * x * (x * x * x - 3.0 * x + 2.0 * c.a)) * br;
}

static double tmax_f_over_r(const CamConst& c) {
    double lo = std::max(c.r_in, c.pt_x0 * c.pt_x0), hi = c.r_out;
    if (hi <= lo) return 0.0;
    const int n = 4096;
    double best = -1, rb = lo;
    for (int i = 0; i <= n; ++i) {
        double r = lo + (hi - lo) * i / n;
        double v = pt_f_host(c, r) / r;
        if (v > best) { best = v; rb = r; }
    }
    double h = (hi - lo) / n;
    double A = std::max(rb - h, lo), B = std::min(rb + h, hi);
    const double gr = 0.6180339887498949;
    for (int it = 0; it < 200; ++it) {
        double cc = B - gr * (B - A), d = A + gr * (B - A);
        if (pt_f_host(c, cc) / cc > pt_f_host(c, d) / d) B = d; else A = cc;
    }
    double rr = 0.5 * (A + B);
    return std::max(best, pt_f_host(c, rr) / rr);
}

// Covariant metric at a point, FP64 host copy of metric_cov (for the observer normalisation).
static double gtt_at(const SiriusHipMetric& m, const double x[4]) {
    if (m.family == SIRIUS_METRIC_MORRIS_THORNE) return -1.0;
    if (m.family == SIRIUS_METRIC_ALCUBIERRE) {
        double rs = std::sqrt(x[1] * x[1] + x[2] * x[2] + x[3] * x[3]);
        double f = (std::tanh(m.warp_sigma * (rs + m.warp_R)) - std::tanh(m.warp_sigma * (rs - m.warp_R))) /
                   (2.0 * std::tanh(m.warp_sigma * m.warp_R));
        return -(1.0 - m.warp_v * m.warp_v * f * f);
    }
    double a2 = m.a * m.a;
    double R2 = x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
    double w = R2 - a2;
    double r2 = 0.5 * (w + std::sqrt(std::max(w * w + 4 * a2 * x[3] * x[3], 1e-20)));
    r2 = std::max(r2, 1e-10);
    double r = std::sqrt(r2);
    double H = (2 * m.M * r - m.Q * m.Q) * r2 / (r2 * r2 + a2 * x[3] * x[3]);
    return -1.0 + H - m.Lambda * R2 / 3.0;
}

bool make_cam_const(const SiriusHipCamera& cam, const SiriusHipMetric& m, int64_t offset, CamConst& c,
                    std::string& err) {
    std::memset(&c, 0, sizeof(c));
    if (cam.width <= 0 || cam.height <= 0) { err = "camera width/height must be positive"; return false; }
    if (!(cam.fov_deg > 0 && cam.fov_deg < 180)) { err = "camera fov_deg must be in (0, 180)"; return false; }
    if (!(cam.r_obs > 0)) { err = "camera r_obs must be positive"; return false; }
    c.r_obs = cam.r_obs;
    c.theta_obs = std::clamp(cam.theta_obs, 0.01, M_PI - 0.01);  // GTRC001A.cpp:25
    c.phi_obs = cam.phi_obs;
    c.tan_half = std::tan(cam.fov_deg * M_PI / 360.0);
    c.aspect = (double)cam.width / cam.height;
    c.cyaw = std::cos(cam.yaw); c.syaw = std::sin(cam.yaw);
    c.cpitch = std::cos(cam.pitch); c.spitch = std::sin(cam.pitch);
    c.croll = std::cos(cam.roll); c.sroll = std::sin(cam.roll);
    double st = std::sin(c.theta_obs), ct = std::cos(c.theta_obs), sp = std::sin(c.phi_obs), cp = std::cos(c.phi_obs);
    c.er[0] = st * cp; c.er[1] = st * sp; c.er[2] = ct;
    c.et[0] = ct * cp; c.et[1] = ct * sp; c.et[2] = -st;
    c.ep[0] = -sp; c.ep[1] = cp; c.ep[2] = 0;
    c.width = cam.width;
    c.height = cam.height;
    c.family = m.family;
    c.offset = offset;
    c.M = m.M; c.a = m.a; c.Q2 = m.Q * m.Q; c.Lambda = m.Lambda;
    c.b0 = m.throat_b0;
    c.wv = m.warp_v; c.wR = m.warp_R; c.wsig = m.warp_sigma;
    if (m.family == SIRIUS_METRIC_ALCUBIERRE) {
        if (!(m.warp_sigma > 0 && m.warp_R > 0)) { err = "warp sigma and R must be positive"; return false; }
        // the shape function's exponential stays inside the FP32 range for sigma R <= 40
        if (m.warp_sigma * m.warp_R > 40.0) { err = "warp sigma R must be at most 40"; return false; }
        c.winv2tanh = 1.0 / (2.0 * std::tanh(m.warp_sigma * m.warp_R));
        c.wE = std::exp(-4.0 * m.warp_sigma * m.warp_R);
        c.wN = (1.0 - c.wE) / std::tanh(m.warp_sigma * m.warp_R);
    }
    if (m.family == SIRIUS_METRIC_MORRIS_THORNE) {
        if (!(m.throat_b0 > 0) || cam.r_obs <= m.throat_b0) { err = "wormhole needs 0 < b0 < r_obs"; return false; }
        c.l_obs = std::sqrt(cam.r_obs * cam.r_obs - m.throat_b0 * m.throat_b0);
    }
    for (double v : {m.M, m.a, m.Q, m.Lambda, m.throat_b0, m.warp_v, m.warp_R, m.warp_sigma, cam.r_obs, cam.theta_obs,
                     cam.phi_obs, cam.disk_r_in, cam.disk_r_out})
        if (!std::isfinite(v)) { err = "camera and metric parameters must be finite"; return false; }
    if (m.family == SIRIUS_METRIC_KERR_SCHILD) {
        double d = m.M * m.M - m.a * m.a - m.Q * m.Q;
        if (!(m.M > 0) || d < 0) { err = "Kerr-Schild family needs M > 0 and M^2 >= a^2 + Q^2"; return false; }
        if (std::fabs(m.M - 1.0) > 1e-12) { err = "units: set M = 1 (G = c = M = 1)"; return false; }
        double rplus = m.M + std::sqrt(d);
        c.r_cap = 0.0;  // set from opts by the caller
        c.r_in = cam.disk_r_in > 0 ? cam.disk_r_in : isco(m.a);
        c.r_out = cam.disk_r_out;
        c.r_ph_retro = photon_orbit(std::fabs(m.a), false);
        // Page-Thorne constants (M = 1)
        double a = m.a;
        c.pt_x0 = std::sqrt(isco(a));
        double ac = std::acos(std::clamp(a, -1.0, 1.0));
        c.pt_x[0] = 2.0 * std::cos((ac - M_PI) / 3.0);
        c.pt_x[1] = 2.0 * std::cos((ac + M_PI) / 3.0);
        c.pt_x[2] = -2.0 * std::cos(ac / 3.0);
        for (int i = 0; i < 3; ++i) {
            double xi = c.pt_x[i], xj = c.pt_x[(i + 1) % 3], xk = c.pt_x[(i + 2) % 3];
            c.pt_c[i] = 3.0 * (xi - a) * (xi - a) / (xi * (xi - xj) * (xi - xk));
        }
        static ConstCache<5, double> fmax_cache;
        const double key[5] = {c.M, c.a, c.Q2, c.r_in, c.r_out};
        c.fmax = fmax_cache.get(key, [&] { return tmax_f_over_r(c); });
        c.r_cap = rplus;  // multiplied by horizon_factor in trace_batch
    }
    double x[4] = {0, c.r_obs * c.er[0], c.r_obs * c.er[1], c.r_obs * c.er[2]};
    double gtt = gtt_at(m, x);
    if (!(gtt < 0)) { err = "observer is not static (g_tt >= 0 at the camera)"; return false; }
    c.obs_norm = 1.0 / std::sqrt(-gtt);
    // the metric at every ray's initial position, for the devices' initial states
    {
        double xo[4] = {0.0, 0.0, 0.0, 0.0};
        if (c.family == 1) xo[1] = c.l_obs;
        else
            for (int i = 0; i < 3; ++i) xo[i + 1] = c.r_obs * c.er[i];
        double g[4][4];
        metric_cov(make_spacetime<double>(c), xo, g);
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = 0; nu < 4; ++nu) c.g_obs[4 * mu + nu] = g[mu][nu];
    }
    return true;
}

}  // namespace sirius

using namespace sirius;

// ---------------------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------------------
namespace {

const char* status_name(int rc) {
    switch (rc) {
        case SIRIUS_HIP_OK: return "ok";
        case SIRIUS_HIP_E_INVALID: return "invalid argument";
        case SIRIUS_HIP_E_DEVICE: return "device error";
        case SIRIUS_HIP_E_NOMEM: return "out of device memory";
        case SIRIUS_HIP_E_STATE: return "bad state";
        default: return "unsupported";
    }
}

}  // namespace

Handle::~Handle() { release(); }

void Handle::release() {
    for (auto& kv : bufs) if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    bufs.clear();
    cam_dev_held = crit_dev_held = nullptr;
    for (int i = 0; i < 2; ++i) {
        if (star_dev[i]) (void)hipFree(star_dev[i]);
        star_dev[i] = nullptr;
        star_n[i] = 0;
    }
    if (fb_dev) (void)hipFree(fb_dev);
    fb_dev = nullptr;
    if (fb_host) (void)hipHostFree(fb_host);
    fb_host = nullptr;
    fb_n = 0;
    if (bg_dev) (void)hipFree(bg_dev);
    bg_dev = nullptr;
    if (pinned) (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_bytes = 0;
    image.clear();
    image.shrink_to_fit();
    mlp_release(*this);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev2) (void)hipEventDestroy(ev2);
    if (ev3) (void)hipEventDestroy(ev3);
    if (ev4) (void)hipEventDestroy(ev4);
    ev0 = ev1 = ev2 = ev3 = ev4 = nullptr;
    delete last;
    last = nullptr;
    for (hipEvent_t* ev : {&ev_gate, &ev_band0, &ev_band1, &ev_s0, &ev_s1, &ev_a0, &ev_a1, &ev_p0, &ev_p1}) {
        if (*ev) (void)hipEventDestroy(*ev);
        *ev = nullptr;
    }
    if (stream2) (void)hipStreamDestroy(stream2);
    stream2 = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    ready = false;
}

int Handle::fail(int rc, const std::string& msg) {
    last_error = std::string(status_name(rc)) + ": " + msg;
    return rc;
}

int Handle::hip_fail(hipError_t e, const char* what) {
    last_error = std::string("device error: ") + what + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? SIRIUS_HIP_E_NOMEM : SIRIUS_HIP_E_DEVICE;
}

int Handle::ensure_device() {
    if (ready) return SIRIUS_HIP_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(SIRIUS_HIP_E_DEVICE, "no HIP device visible");
    e = hipSetDevice(0);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, 0);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SIRIUS_HIP_E_DEVICE, std::string("built for gfx950, device is ") + prop.gcnArchName);
    cu_count = prop.multiProcessorCount;
    max_threads = prop.maxThreadsPerBlock;
    device_name = prop.name;
    total_mem = prop.totalGlobalMem;
    if ((e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "stream");
    if ((e = hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking)) != hipSuccess) return hip_fail(e, "stream");
    for (hipEvent_t* ev : {&ev_gate, &ev_band0, &ev_band1, &ev_s0, &ev_s1, &ev_a0, &ev_a1, &ev_p0, &ev_p1})
        if ((e = hipEventCreate(ev)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&ev0)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&ev1)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&ev2)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&ev3)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&ev4)) != hipSuccess) return hip_fail(e, "event");
    ready = true;
    return SIRIUS_HIP_OK;
}

void* Handle::scratch(const char* name, size_t bytes, int* rc) {
    if (arena) {
        // no-allocation mode: the same name within one call is the same sub-buffer
        for (int i = 0; i < arena_n; ++i)
            if (std::strcmp(arena_slots[i].name, name) == 0) {
                if (arena_slots[i].bytes >= bytes) return arena + arena_slots[i].off;
                *rc = fail(SIRIUS_HIP_E_INVALID, std::string("no-allocation scratch: buffer ") + name + " grew");
                return nullptr;
            }
        size_t off = (arena_used + 255) & ~(size_t)255;
        if (arena_n >= kArenaSlots || off + bytes > arena_cap) {
            *rc = fail(SIRIUS_HIP_E_INVALID, "batch exceeds the no-allocation scratch (" +
                                                 std::to_string(arena_cap >> 20) + " MiB)");
            return nullptr;
        }
        arena_slots[arena_n++] = {name, off, bytes};
        arena_used = off + bytes;
        return arena + off;
    }
    Buf& b = bufs[name];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    // a replaced buffer no longer holds an upload (a new allocation may reuse its address)
    if (b.ptr == cam_dev_held) cam_dev_held = nullptr;
    if (b.ptr == crit_dev_held) crit_dev_held = nullptr;
    // nor the last trace's results: fetching them afterwards is an error, not a read of freed memory
    if (b.ptr && last) {
        const void* held[] = {last->cls,      last->flags,    last->theta, last->phi,       last->lam,
                              last->disk_r,   last->redshift, last->null_drift, last->steps, last->rejected,
                              last->landings, last->intensity, last->state, last->vol_acc};
        for (const void* p : held)
            if (p == b.ptr) {
                delete last;
                last = nullptr;
                break;
            }
    }
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.ptr, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        b.ptr = nullptr;
        *rc = hip_fail(e, "hipMalloc");
        return nullptr;
    }
    b.bytes = bytes;
    return b.ptr;
}

// ---------------------------------------------------------------------------------------
// Trace pipeline
// ---------------------------------------------------------------------------------------
namespace sirius {

static int validate_opts(Handle& H, const SiriusHipTraceOpts& o) {
    if (o.n_bins < 0 || o.n_bins > 64) return H.fail(SIRIUS_HIP_E_INVALID, "n_bins must be in [0, 64]");
    if (!(o.rtol > 0) || !(o.atol > 0)) return H.fail(SIRIUS_HIP_E_INVALID, "rtol and atol must be positive");
    if (!(o.h_init > 0) || !(o.h_min > 0) || !(o.h_max >= o.h_min))
        return H.fail(SIRIUS_HIP_E_INVALID, "need 0 < h_min <= h_max and h_init > 0");
    if (o.max_steps <= 0) return H.fail(SIRIUS_HIP_E_INVALID, "max_steps must be positive");
    if (!(o.step_radial_exp >= 0 && o.step_radial_exp <= 8))
        return H.fail(SIRIUS_HIP_E_INVALID, "step_radial_exp must be in [0, 8]");
    if (!(o.r_escape > 0) || !(o.horizon_factor >= 1.0))
        return H.fail(SIRIUS_HIP_E_INVALID, "need r_escape > 0 and horizon_factor >= 1");
    if (!(o.x_min > 0 && o.x_max > o.x_min)) return H.fail(SIRIUS_HIP_E_INVALID, "need 0 < x_min < x_max");
    if (o.integrator != SIRIUS_INTEGRATOR_DP45 && o.integrator != SIRIUS_INTEGRATOR_SYMPLECTIC)
        return H.fail(SIRIUS_HIP_E_INVALID, "integrator");
    if (o.precision != SIRIUS_PRECISION_FP64 && o.precision != SIRIUS_PRECISION_FP32_ESCALATE)
        return H.fail(SIRIUS_HIP_E_INVALID, "precision");
    if (o.emission < 0 || o.emission > 2) return H.fail(SIRIUS_HIP_E_INVALID, "emission");
    if (o.schedule != SIRIUS_SCHEDULE_GRID && o.schedule != SIRIUS_SCHEDULE_PERSISTENT)
        return H.fail(SIRIUS_HIP_E_INVALID, "schedule");
    if (o.block_size != 0 && (o.block_size < 64 || o.block_size > 1024 || o.block_size % 64))
        return H.fail(SIRIUS_HIP_E_INVALID, "block_size must be 0 or a multiple of 64 in [64, 1024]");
    if (o.integrator == SIRIUS_INTEGRATOR_SYMPLECTIC && !(o.symplectic_h > 0))
        return H.fail(SIRIUS_HIP_E_INVALID, "symplectic_h must be positive");
    if (o.use_surrogate && (o.precision != SIRIUS_PRECISION_FP32_ESCALATE || o.schedule != SIRIUS_SCHEDULE_PERSISTENT ||
                            o.integrator != SIRIUS_INTEGRATOR_DP45))
        return H.fail(SIRIUS_HIP_E_UNSUPPORTED, "the surrogate runs on the FP32 persistent DP45 path");
    if (o.use_surrogate && !(o.audit_fraction >= 0 && o.audit_fraction <= 1))
        return H.fail(SIRIUS_HIP_E_INVALID, "audit_fraction must be in [0, 1]");
    if (o.use_surrogate && !(o.surrogate_err_max >= 0 && o.surrogate_margin >= 0 && o.surrogate_band >= 0))
        return H.fail(SIRIUS_HIP_E_INVALID, "surrogate_err_max, surrogate_margin and surrogate_band must be >= 0");
    if (!(o.precision_audit >= 0 && o.precision_audit <= 1))
        return H.fail(SIRIUS_HIP_E_INVALID, "precision_audit must be in [0, 1]");
    return SIRIUS_HIP_OK;
}

int prepare_batch(Handle& H, const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics, int n_metric,
                  const SiriusHipTraceOpts& o, Batch& B) {
    if (!cams || n_cam <= 0 || !metrics || n_metric <= 0) return H.fail(SIRIUS_HIP_E_INVALID, "cameras/metrics");
    int rc = validate_opts(H, o);
    if (rc) return rc;
    B.cc.resize(n_cam);
    int64_t off = 0;
    B.family = -1;
    for (int i = 0; i < n_cam; ++i) {
        if (cams[i].metric_index < 0 || cams[i].metric_index >= n_metric)
            return H.fail(SIRIUS_HIP_E_INVALID, "camera metric_index out of range");
        const SiriusHipMetric& m = metrics[cams[i].metric_index];
        if (m.family < 0 || m.family > 2) return H.fail(SIRIUS_HIP_E_INVALID, "metric family");
        if (B.family < 0) B.family = m.family;
        if (B.family != m.family) return H.fail(SIRIUS_HIP_E_UNSUPPORTED, "one metric family per batch");
        std::string err;
        if (!make_cam_const(cams[i], m, off, B.cc[i], err)) return H.fail(SIRIUS_HIP_E_INVALID, err);
        B.cc[i].r_cap *= o.horizon_factor;
        off += (int64_t)cams[i].width * cams[i].height;
    }
    if (off <= 0 || off > (int64_t)INT32_MAX - (1 << 20)) return H.fail(SIRIUS_HIP_E_INVALID, "ray count out of range");
    B.n = off;
    TraceParams& tp = B.tp;
    std::memset(&tp, 0, sizeof(tp));
    tp.rtol = o.rtol; tp.atol = o.atol; tp.h_init = o.h_init; tp.h_min = o.h_min; tp.h_max = o.h_max;
    tp.far_slope = o.far_slope;
    tp.r_escape = o.r_escape;
    tp.x_min = o.x_min; tp.x_max = o.x_max;
    tp.escalate_margin = o.escalate_margin;
    tp.sym_h = o.symplectic_h;
    tp.vol_j0 = o.vol_j0; tp.vol_h_over_r = o.vol_h_over_r; tp.vol_kappa = o.vol_kappa;
    tp.vol_r_min = o.vol_r_min; tp.vol_r_max = o.vol_r_max; tp.vol_tau_max = o.vol_tau_max;
    tp.esc_edge = o.escalate_disk_edge;
    tp.step_rexp = o.step_radial_exp;
    tp.gate_flags = 0;
    tp.spectrum32 = o.precision == SIRIUS_PRECISION_FP32_ESCALATE ? 1 : 0;
    tp.first_list = nullptr;
    tp.first_count = nullptr;
    tp.max_steps = o.max_steps;
    tp.n_bins = o.n_bins;
    bin_centres(o.x_min, o.x_max, o.n_bins, tp.bin_x0, tp.bin_q);
    tp.emission = B.family == SIRIUS_METRIC_KERR_SCHILD ? o.emission : SIRIUS_EMISSION_NONE;
    tp.sym_order = o.symplectic_order;
    tp.n_cam = n_cam;
    // escalation machinery of the FP32 path: always on for the wormhole throat and the warp wall;
    // for Kerr-Schild on when any of its criteria is (a zero criterion never fires)
    tp.escalate = (o.precision == SIRIUS_PRECISION_FP32_ESCALATE &&
                   (B.family != SIRIUS_METRIC_KERR_SCHILD || o.escalate_margin > 0 || o.escalate_band > 0 ||
                    o.escalate_disk_edge > 0)) ? 1 : 0;
    tp.n_rays = off;
    if (tp.emission == SIRIUS_EMISSION_VOLUMETRIC && !(o.vol_h_over_r > 0 && o.vol_r_max > o.vol_r_min))
        return H.fail(SIRIUS_HIP_E_INVALID, "volumetric parameters");
    void* d = H.scratch("cams", sizeof(CamConst) * n_cam, &rc);
    if (!d) return rc;
    B.cams_dev = (CamConst*)d;
    // every batch ends with a synchronisation, so the device copy of the last batch's cameras is
    // intact: an identical batch (a repeated frame) needs no upload
    if (H.cam_dev_held == d && H.cam_host.size() == B.cc.size() &&
        std::memcmp(H.cam_host.data(), B.cc.data(), sizeof(CamConst) * n_cam) == 0)
        return SIRIUS_HIP_OK;
    H.cam_host = B.cc;  // outlives the asynchronous copy
    H.cam_dev_held = nullptr;
    hipError_t e = hipMemcpyAsync(d, H.cam_host.data(), sizeof(CamConst) * n_cam, hipMemcpyHostToDevice, H.stream);
    if (e != hipSuccess) return H.hip_fail(e, "upload cameras");
    H.cam_dev_held = d;
    return SIRIUS_HIP_OK;
}

// Device output arrays: the caller's (on_device) or library scratch.
int bind_outputs(Handle& H, const SiriusHipRayOutput* out, int64_t n, int n_bins, DevOut& o, bool volumetric) {
    int rc = SIRIUS_HIP_OK;
    auto get = [&](const char* name, void* user, size_t bytes) -> void* {
        if (out && out->on_device && user) return user;
        return H.scratch(name, bytes, &rc);
    };
    o.n_rays = n;
    o.cls = (uint8_t*)get("o_cls", out ? out->cls : nullptr, n);
    o.flags = (uint8_t*)get("o_flags", out ? out->flags : nullptr, n);
    o.theta = (float*)get("o_theta", out ? out->theta : nullptr, 4 * n);
    o.phi = (float*)get("o_phi", out ? out->phi : nullptr, 4 * n);
    o.lam = (float*)get("o_lam", out ? out->lambda : nullptr, 4 * n);
    o.disk_r = (float*)get("o_disk_r", out ? out->disk_r : nullptr, 4 * n);
    o.redshift = (float*)get("o_redshift", out ? out->redshift : nullptr, 4 * n);
    o.null_drift = (float*)get("o_null", out ? out->null_drift : nullptr, 4 * n);
    o.steps = (uint32_t*)get("o_steps", out ? out->steps : nullptr, 4 * n);
    o.rejected = (uint32_t*)H.scratch("o_rej", 4 * n, &rc);
    o.landings = (uint32_t*)H.scratch("o_land", 4 * n, &rc);
    o.intensity = (float*)get("o_int", out ? out->intensity : nullptr, 4 * n * std::max(n_bins, 1));
    // the final state is always kept: the finish kernel derives every output from it
    o.state = (double*)get("o_state", out ? out->state : nullptr, 64 * n);
    // volumetric emission accumulates in FP64 (the finish kernel rounds it into intensity)
    o.vol_acc = volumetric ? (double*)H.scratch("o_vacc", 8 * n * std::max(n_bins, 1), &rc) : nullptr;
    // the arrays about to be written no longer hold the last results: callers that report
    // their rays as the last results (trace_batch, launch) record them again afterwards
    if (H.last) {
        const void* mine[] = {o.cls, o.flags, o.theta, o.phi, o.lam, o.disk_r, o.redshift, o.null_drift,
                              o.steps, o.intensity, o.state};
        const void* held[] = {H.last->cls, H.last->flags, H.last->theta, H.last->phi, H.last->lam,
                              H.last->disk_r, H.last->redshift, H.last->null_drift, H.last->steps,
                              H.last->intensity, H.last->state};
        bool shared = false;
        for (const void* a : mine)
            for (const void* b : held) shared |= a && a == b;
        if (shared) {
            delete H.last;
            H.last = nullptr;
        }
    }
    return rc;
}

// Library-owned result arrays of the surrogate audit (separate from the batch's outputs).
static int bind_audit_outputs(Handle& H, int64_t n, int n_bins, DevOut& o) {
    int rc = SIRIUS_HIP_OK;
    o.n_rays = n;
    o.cls = (uint8_t*)H.scratch("a_cls", n, &rc);
    o.flags = (uint8_t*)H.scratch("a_flags", n, &rc);
    o.theta = (float*)H.scratch("a_theta", 4 * n, &rc);
    o.phi = (float*)H.scratch("a_phi", 4 * n, &rc);
    o.lam = (float*)H.scratch("a_lam", 4 * n, &rc);
    o.disk_r = (float*)H.scratch("a_disk_r", 4 * n, &rc);
    o.redshift = (float*)H.scratch("a_redshift", 4 * n, &rc);
    o.null_drift = (float*)H.scratch("a_null", 4 * n, &rc);
    o.steps = (uint32_t*)H.scratch("a_steps", 4 * n, &rc);
    o.rejected = (uint32_t*)H.scratch("a_rej", 4 * n, &rc);
    o.landings = (uint32_t*)H.scratch("a_land", 4 * n, &rc);
    o.intensity = (float*)H.scratch("a_int", 4 * n * std::max(n_bins, 1), &rc);
    o.state = (double*)H.scratch("a_state", 64 * n, &rc);
    o.vol_acc = nullptr;  // the audit serves thin-disk batches only
    return rc;
}

int copy_outputs_to_host(Handle& H, const SiriusHipRayOutput* out, const DevOut& o, int n_bins) {
    if (!out || out->on_device) return SIRIUS_HIP_OK;
    int64_t n = o.n_rays;
    struct { void* dst; const void* src; size_t bytes; } c[] = {
        {out->cls, o.cls, (size_t)n},           {out->flags, o.flags, (size_t)n},
        {out->theta, o.theta, 4 * (size_t)n},   {out->phi, o.phi, 4 * (size_t)n},
        {out->lambda, o.lam, 4 * (size_t)n},    {out->disk_r, o.disk_r, 4 * (size_t)n},
        {out->redshift, o.redshift, 4 * (size_t)n}, {out->null_drift, o.null_drift, 4 * (size_t)n},
        {out->steps, o.steps, 4 * (size_t)n},   {out->intensity, o.intensity, 4 * (size_t)n * n_bins},
        {out->state, o.state, 64 * (size_t)n},
    };
    for (auto& x : c) {
        if (!x.dst || !x.src || !x.bytes) continue;
        hipError_t e = hipMemcpyAsync(x.dst, x.src, x.bytes, hipMemcpyDeviceToHost, H.stream);
        if (e != hipSuccess) return H.hip_fail(e, "copy results to host");
    }
    hipError_t e = hipStreamSynchronize(H.stream);
    if (e != hipSuccess) return H.hip_fail(e, "synchronize");
    return SIRIUS_HIP_OK;
}

// Per-family algorithmic flops of one DP45 attempt (6 RHS + 426, fixed constants).
double flops_per_step(int family) {
    switch (family) {
        case 1: return 576.0;
        case 2: return 678.0;
        default: return 1326.0;
    }
}

// Algorithmic flops of one fixed step of the symplectic integrator: per Strang substep of Tao's
// extended phase space, 4 right-hand-side evaluations, 4 half-step updates of 7 components
// (56) and the rotation of the 4 coordinate pairs (72); 3 substeps for Yoshida-4, 1 for order 2.
double flops_per_symplectic_step(int family, int order) {
    double rhs = family == 1 ? 25.0 : (family == 2 ? 42.0 : 150.0);
    return (order >= 4 ? 3.0 : 1.0) * (4.0 * rhs + 56.0 + 72.0);
}

int run_trace(Handle& H, Batch& B, const SiriusHipTraceOpts& o, DevOut& dout, const int32_t* list_dev,
              const uint32_t* list_count_dev, int64_t list_cap, SiriusHipStats* stats, bool time_it) {
    hipError_t e;
    int rc = SIRIUS_HIP_OK;
    bool fp64 = o.precision == SIRIUS_PRECISION_FP64;
    int block = o.block_size > 0 ? o.block_size : 256;
    int64_t n_work = list_dev ? list_cap : B.n;
    bool persistent = o.schedule == SIRIUS_SCHEDULE_PERSISTENT && o.integrator == SIRIUS_INTEGRATOR_DP45;
    // work[0] bulk queue head, work[1] escalation count, work[2..3] bulk attempts (u64),
    // work[4] escalation queue head, work[5] unused, work[6..7] escalation attempts (u64),
    // work[8..10] surrogate list counts (A, B, audit), work[11] audit queue head,
    // work[12] band queue head, work[13] unused, work[14..15] band attempts (u64),
    // work[16..18] gate counts (A, B, L), work[19] surrogate candidates, work[20] candidates the
    // calibrated error gate refused
    uint32_t* work = (uint32_t*)H.scratch("work", 128, &rc);
    if (rc) return rc;
    unsigned int* audit_acc = nullptr;  // surrogate audit: largest direction error, class mismatches
    if ((e = launch_fill(work, 32, 0u, H.stream)) != hipSuccess) return H.hip_fail(e, "fill");
    H.last_band = false;
    H.last_audit = false;  // set again below only when this call audits surrogate answers
    if (time_it) (void)hipEventRecord(H.ev0, H.stream);
    if (o.integrator == SIRIUS_INTEGRATOR_SYMPLECTIC) {
        if (list_dev) return H.fail(SIRIUS_HIP_E_UNSUPPORTED, "symplectic integrator with a ray list");
        if ((e = launch_symplectic(fp64, B.family, B.cams_dev, B.tp, dout, block, H.stream)) != hipSuccess)
            return H.hip_fail(e, "symplectic launch");
    } else if (!persistent) {
        // Grid schedule: one lane per ray, run to completion, no escalation queue (FP32 lanes
        // run to the end in FP32). It is the tracer kernel with a static assignment: a grid of
        // one wave per 64 rays, each wave taking one batch of 64 consecutive list slots when all
        // its lanes are idle, so both schedules run the same compiled code, bit for bit.
        TraceParams tp = B.tp;
        tp.escalate = 0;
        int grid = (int)std::max<int64_t>(1, (n_work + block - 1) / block);
        if ((e = launch_trace_persistent(fp64, B.family, B.tp.emission, B.tp.n_cam > 1, B.cams_dev, tp, dout, list_dev,
                                         list_count_dev, n_work, work, work + 1, nullptr,
                                         (unsigned long long*)(work + 2), grid, block, 64, 0, H.stream)) !=
            hipSuccess)
            return H.hip_fail(e, "grid launch");
    } else if (fp64) {
        TraceParams tp = B.tp;
        tp.escalate = 0;
        int bpc = o.blocks_per_cu > 0 ? o.blocks_per_cu
                                      : persistent_blocks_per_cu(true, B.family, B.tp.emission, B.tp.n_cam > 1, block);
        int refill = o.refill_min > 0 ? std::min(o.refill_min, 64) : 16;
        if ((e = launch_trace_persistent(true, B.family, B.tp.emission, B.tp.n_cam > 1, B.cams_dev, tp, dout, list_dev,
                                         list_count_dev, n_work, work, work + 1, nullptr,
                                         (unsigned long long*)(work + 2), std::max(1, bpc * H.cu_count), block, refill,
                                         0, H.stream)) != hipSuccess)
            return H.hip_fail(e, "persistent launch");
        if (time_it) (void)hipEventRecord(H.ev1, H.stream);
    } else {
        // FP32 path. (1) Gate (Kerr-Schild): every ray's distance from the critical curve; rays
        // within critical_band are flagged, rays within escalate_band go to list B and are traced
        // in FP64 on the second stream while (2) the FP32 persistent tracer runs list A; (3) FP32
        // rays that meet an in-flight criterion (inner-edge disk crossing, turning point inside
        // the photon region, warp wall) are re-traced in FP64 from their camera pixel after the
        // bulk. Escalated rays keep the path's step caps with the escalation tolerance.
        int32_t* esc = (int32_t*)H.scratch("esc_list", 4 * (size_t)std::max<int64_t>(n_work, 1), &rc);
        if (rc) return rc;
        const int32_t* list_a = list_dev;
        const uint32_t* count_a = list_count_dev;
        int64_t cap_a = n_work;
        int32_t* list_b = nullptr;
        // the surrogate answers before the integrator (Kerr-Schild, thin disk or no emission)
        const bool sur = o.use_surrogate != 0;
        if (sur && (list_dev || B.family != SIRIUS_METRIC_KERR_SCHILD || B.tp.emission == SIRIUS_EMISSION_VOLUMETRIC))
            return H.fail(SIRIUS_HIP_E_UNSUPPORTED, "the surrogate serves whole Kerr-Schild batches without volumetric emission");
        if (sur && (!H.mlp || !H.mlp->ready))
            return H.fail(SIRIUS_HIP_E_STATE, "use_surrogate needs a network (sirius_hip_mlp_init / set_params / surrogate_train)");
        // the gate runs for every whole Kerr-Schild batch: besides the critical flags and the FP64
        // band it orders the queue (rays near the critical curve, the longest, first)
        bool gate = !sur && !list_dev && B.family == SIRIUS_METRIC_KERR_SCHILD;
        bool band = (gate || sur) && o.escalate_band > 0;
        TraceParams t32 = B.tp;
        t32.gate_flags = (gate || sur) ? 1 : 0;
        int32_t* list_audit = nullptr;
        if (sur) {
            int32_t* la = (int32_t*)H.scratch("gate_a", 4 * (size_t)B.n, &rc);
            list_b = (int32_t*)H.scratch("gate_b", 4 * (size_t)B.n, &rc);
            int32_t* ll = (int32_t*)H.scratch("gate_l", 4 * (size_t)B.n, &rc);
            list_audit = (int32_t*)H.scratch("sur_audit", 4 * (size_t)B.n, &rc);
            if (rc) return rc;
            if (time_it) (void)hipEventRecord(H.ev_s0, H.stream);
            if ((rc = surrogate_stage(H, B, o, dout, la, list_b, ll, list_audit, work))) return rc;
            if (time_it) (void)hipEventRecord(H.ev_s1, H.stream);
            list_a = la;
            count_a = work + 16;
            cap_a = B.n;
            t32.first_list = ll;
            t32.first_count = work + 18;
        }
        if (gate) {
            float* tabs = upload_critical_tables(H, B, &rc);
            int32_t* la = (int32_t*)H.scratch("gate_a", 4 * (size_t)B.n, &rc);
            list_b = (int32_t*)H.scratch("gate_b", 4 * (size_t)B.n, &rc);
            int32_t* ll = (int32_t*)H.scratch("gate_l", 4 * (size_t)B.n, &rc);
            if (rc) return rc;
            uint32_t* counts = work + 16;
            const double long_band = SIRIUS_LONG_BAND;  // the longest rays first
            if ((e = launch_gate(B.cams_dev, (int)B.cc.size(), tabs, B.n, std::max(o.escalate_band, 0.0),
                                 std::max(o.critical_band, 0.0), long_band, la, list_b, ll, counts, dout.flags,
                                 H.stream)) != hipSuccess)
                return H.hip_fail(e, "gate launch");
            list_a = la;
            count_a = counts;
            cap_a = B.n;
            t32.first_list = ll;
            t32.first_count = counts + 2;
        }
        // escalated rays: FP64 arithmetic, the path's step caps, the escalation tolerance
        TraceParams t64 = t32;
        t64.escalate = 0;
        t64.first_list = nullptr;  // the FP64 kernels trace only their own lists
        t64.first_count = nullptr;
        // Default tolerances of the escalated rays (the path's step caps kept), from the 1024x1024
        // scenes checked at every pixel against the plain FP64 frame: the warp bubble's wall
        // crossers at 3e-10 (at 1e-8 or 1e-9 the worst missed by 1.2e-4 rad, while the plain FP64
        // frame is itself converged to 1.6e-5 there; at 3e-10 every pixel is within 1.6e-5); the
        // wormhole's near-throat rays at 1e-8 (at the path's 1e-6 the worst missed by 1.2e-3 rad;
        // at 1e-8 by 6.3e-6). Kerr-Schild escalations keep the path's rtol.
        double ertol = o.escalate_rtol;
        if (!(ertol > 0)) {
            if (B.family == SIRIUS_METRIC_ALCUBIERRE) ertol = std::min(o.rtol, 3e-10);
            else if (B.family == SIRIUS_METRIC_MORRIS_THORNE) ertol = std::min(o.rtol, 1e-8);
            else ertol = 0.0;
        }
        if (ertol > 0) {
            t64.rtol = ertol;
            t64.atol = o.atol * (ertol / o.rtol);
        }
        if (band) {
            // FP64 band on the second stream, small blocks so the long rays spread over the chip.
            // The band's length stays on the device (no host round trip): the persistent grid is
            // sized for the whole chip and waves with nothing to trace retire at once.
            (void)hipEventRecord(H.ev_gate, H.stream);
            (void)hipStreamWaitEvent(H.stream2, H.ev_gate, 0);
            int bpc64 = persistent_blocks_per_cu(true, B.family, B.tp.emission, B.tp.n_cam > 1, 64);
            (void)hipEventRecord(H.ev_band0, H.stream2);
            if ((e = launch_trace_persistent(true, B.family, B.tp.emission, B.tp.n_cam > 1, B.cams_dev, t64, dout, list_b,
                                             work + 17, B.n, work + 12, work + 13, nullptr,
                                             (unsigned long long*)(work + 14), std::max(1, bpc64 * H.cu_count), 64, 1,
                                             SIRIUS_FLAG_ESCALATED, H.stream2)) != hipSuccess)
                return H.hip_fail(e, "band launch");
            (void)hipEventRecord(H.ev_band1, H.stream2);
        }
        int bpc = o.blocks_per_cu > 0 ? o.blocks_per_cu
                                      : persistent_blocks_per_cu(false, B.family, B.tp.emission, B.tp.n_cam > 1, block);
        int grid = std::max(1, bpc * H.cu_count);
        int refill = o.refill_min > 0 ? std::min(o.refill_min, 64) : 16;
        if (time_it) (void)hipEventRecord(H.ev0, H.stream);
        if ((e = launch_trace_persistent(false, B.family, B.tp.emission, B.tp.n_cam > 1, B.cams_dev, t32, dout, list_a,
                                         count_a, cap_a, work, work + 1, esc, (unsigned long long*)(work + 2), grid,
                                         block, refill, 0, H.stream)) != hipSuccess)
            return H.hip_fail(e, "persistent launch");
        if (time_it) (void)hipEventRecord(H.ev1, H.stream);
        if (B.tp.escalate) {
            int bpc64 = persistent_blocks_per_cu(true, B.family, B.tp.emission, B.tp.n_cam > 1, 64);
            if ((e = launch_trace_persistent(true, B.family, B.tp.emission, B.tp.n_cam > 1, B.cams_dev, t64, dout, esc,
                                             work + 1, cap_a, work + 4, work + 5, nullptr,
                                             (unsigned long long*)(work + 6), std::max(1, bpc64 * H.cu_count), 64, 1,
                                             SIRIUS_FLAG_ESCALATED, H.stream)) != hipSuccess)
                return H.hip_fail(e, "escalation launch");
        }
        if (band) (void)hipStreamWaitEvent(H.stream, H.ev_band1, 0);
        H.last_band = band;
        H.last_audit = sur && o.audit_fraction > 0;
        if (H.last_audit) {
            // audit after the batch (timed apart): FP64 re-traces of the audited answers
            if (time_it) (void)hipEventRecord(H.ev_a0, H.stream);
            DevOut oa;
            if ((rc = bind_audit_outputs(H, B.n, o.n_bins, oa))) return rc;
            unsigned int* aacc = (unsigned int*)H.scratch("audit_acc", 64, &rc);
            if (rc) return rc;
            audit_acc = aacc;
            if ((e = launch_fill(aacc, 2, 0u, H.stream)) != hipSuccess) return H.hip_fail(e, "fill");
            unsigned long long* aatt = (unsigned long long*)H.scratch("audit_att", 8, &rc);
            if (rc) return rc;
            TraceParams ta = fp64_reference_params(B.tp);
            if ((e = launch_fill((uint32_t*)aatt, 2, 0u, H.stream)) != hipSuccess) return H.hip_fail(e, "fill");
            int bpc64 = persistent_blocks_per_cu(true, B.family, B.tp.emission, B.tp.n_cam > 1, 64);
            if ((e = launch_trace_persistent(true, B.family, B.tp.emission, B.tp.n_cam > 1, B.cams_dev, ta, oa,
                                             list_audit, work + 10, B.n, work + 11, work + 5, nullptr, aatt,
                                             std::max(1, bpc64 * H.cu_count), 64, 1, SIRIUS_FLAG_AUDITED,
                                             H.stream)) != hipSuccess)
                return H.hip_fail(e, "audit launch");
            if ((e = launch_sur_audit(B.cams_dev, ta, oa, dout, list_audit, work + 10, B.n, aacc, H.stream)) !=
                hipSuccess)
                return H.hip_fail(e, "audit compare");
            if (time_it) (void)hipEventRecord(H.ev_a1, H.stream);
        }
    }
    if (time_it) {
        if (!persistent) (void)hipEventRecord(H.ev1, H.stream);
        (void)hipEventRecord(H.ev4, H.stream);
    }
    if ((e = launch_finish(B.family, B.tp.emission, B.cams_dev, B.tp, dout, H.stream)) != hipSuccess)
        return H.hip_fail(e, "finish launch");
    // precision audit of the FP32 path (timed apart): a seeded fraction of the traced rays again by
    // the plain FP64 integrator, compared with the batch's own results
    unsigned int* pacc = nullptr;
    unsigned int* phist = nullptr;
    if (!fp64 && o.precision_audit > 0 && o.integrator == SIRIUS_INTEGRATOR_DP45 && !list_dev) {
        if (B.tp.emission == SIRIUS_EMISSION_VOLUMETRIC)
            return H.fail(SIRIUS_HIP_E_UNSUPPORTED, "the precision audit serves thin-disk and emission-free batches");
        if (time_it) (void)hipEventRecord(H.ev_p0, H.stream);
        DevOut oa;
        if ((rc = bind_audit_outputs(H, B.n, o.n_bins, oa))) return rc;
        int32_t* plist = (int32_t*)H.scratch("prec_list", 4 * (size_t)B.n, &rc);
        pacc = (unsigned int*)H.scratch("prec_acc", 4 * 8, &rc);
        phist = (unsigned int*)H.scratch("prec_hist", 4 * 128, &rc);
        unsigned long long* patt = (unsigned long long*)H.scratch("prec_att", 8, &rc);
        if (rc) return rc;
        // work[21] audit list length, work[22] its queue head, work[23] the FP64 kernel's escalation count
        if ((e = launch_fill(pacc, 8, 0u, H.stream)) != hipSuccess || (e = launch_fill(phist, 128, 0u, H.stream)) != hipSuccess ||
            (e = launch_fill((uint32_t*)patt, 2, 0u, H.stream)) != hipSuccess ||
            (e = launch_fill(work + 21, 3, 0u, H.stream)) != hipSuccess)
            return H.hip_fail(e, "fill");
        TraceParams tr = fp64_reference_params(B.tp);
        int bpc64 = persistent_blocks_per_cu(true, B.family, B.tp.emission, B.tp.n_cam > 1, 64);
        if ((e = launch_precision_select(dout, B.n, o.precision_audit, o.audit_seed, plist, work + 21, H.stream)) !=
                hipSuccess ||
            (e = launch_trace_persistent(true, B.family, B.tp.emission, B.tp.n_cam > 1, B.cams_dev, tr, oa, plist,
                                         work + 21, B.n, work + 22, work + 23, nullptr, patt,
                                         std::max(1, bpc64 * H.cu_count), 64, 1, SIRIUS_FLAG_AUDITED, H.stream)) !=
                hipSuccess ||
            (e = launch_precision_compare(B.cams_dev, tr, oa, dout, plist, work + 21, B.n, pacc, phist, H.stream)) !=
                hipSuccess)
            return H.hip_fail(e, "precision audit");
        if (time_it) (void)hipEventRecord(H.ev_p1, H.stream);
    }
    if (stats) {
        unsigned long long* acc = (unsigned long long*)H.scratch("stat_acc", 8 * 8, &rc);
        if (rc) return rc;
        uint32_t* hist = (uint32_t*)H.scratch("stat_hist", 4 * SIRIUS_DRIFT_BINS, &rc);
        if (rc) return rc;
        if ((e = launch_fill((uint32_t*)acc, 16, 0u, H.stream)) != hipSuccess) return H.hip_fail(e, "fill");
        if ((e = launch_fill(hist, SIRIUS_DRIFT_BINS, 0u, H.stream)) != hipSuccess) return H.hip_fail(e, "fill");
        if ((e = launch_reduce_stats(dout, B.n, acc, hist, H.stream)) != hipSuccess) return H.hip_fail(e, "reduce");
        unsigned long long h[8] = {0};
        if ((e = hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, H.stream)) != hipSuccess)
            return H.hip_fail(e, "stats copy");
        uint32_t hh[SIRIUS_DRIFT_BINS];
        if ((e = hipMemcpyAsync(hh, hist, sizeof(hh), hipMemcpyDeviceToHost, H.stream)) != hipSuccess)
            return H.hip_fail(e, "stats copy");
        uint32_t wk[32] = {0};
        if (persistent) {
            if ((e = hipMemcpyAsync(wk, work, sizeof(wk), hipMemcpyDeviceToHost, H.stream)) != hipSuccess)
                return H.hip_fail(e, "stats copy");
        }
        if ((e = hipStreamSynchronize(H.stream)) != hipSuccess) return H.hip_fail(e, "synchronize");
        // h: attempts, rejected, landings, drift_max bits, escalated, surrogate, audited, esc attempts
        stats->rays += (uint64_t)B.n;
        stats->steps_accepted += h[0] - h[1];
        stats->steps_rejected += h[1];
        stats->steps_landing += h[2];
        double dm;
        std::memcpy(&dm, &h[3], sizeof(dm));
        stats->null_drift_max = std::max(stats->null_drift_max, dm);
        {
            // median from the log10 histogram: the centre of the bin holding the middle ray
            uint64_t total = 0, run = 0;
            for (int b = 0; b < SIRIUS_DRIFT_BINS; ++b) total += hh[b];
            int mb = 0;
            for (; mb < SIRIUS_DRIFT_BINS; ++mb) {
                run += hh[mb];
                if (2 * run >= total) break;
            }
            stats->null_drift_median = (total == 0 || mb == 0) ? 0.0 : std::pow(10.0, (mb + 0.5) / 16.0 - 20.0);
        }
        stats->rays_escalated += h[4];
        stats->rays_audited += h[6];
        unsigned long long bulk = h[0];
        if (persistent) {
            unsigned long long esc_att = 0, band_att = 0;
            std::memcpy(&bulk, &wk[2], sizeof(bulk));
            std::memcpy(&esc_att, &wk[6], sizeof(esc_att));
            std::memcpy(&band_att, &wk[14], sizeof(band_att));
            stats->steps_escalated += esc_att + band_att;
            stats->flops_escalated += (double)(esc_att + band_att) * flops_per_step(B.family);
        }
        stats->steps_bulk += bulk;
        stats->flops += (double)bulk * (o.integrator == SIRIUS_INTEGRATOR_SYMPLECTIC
                                            ? flops_per_symplectic_step(B.family, o.symplectic_order)
                                            : flops_per_step(B.family));
        if (H.last_audit && audit_acc) {
            unsigned int aa[2] = {0, 0};
            if ((e = hipMemcpyAsync(aa, audit_acc, sizeof(aa), hipMemcpyDeviceToHost, H.stream)) != hipSuccess ||
                (e = hipStreamSynchronize(H.stream)) != hipSuccess)
                return H.hip_fail(e, "audit copy");
            float err;
            std::memcpy(&err, &aa[0], 4);
            stats->audit_dir_err_max = std::max(stats->audit_dir_err_max, (double)err);
            stats->audit_class_mismatch += aa[1];
        }
        if (pacc) {
            unsigned int pa[4], ph[128];
            if ((e = hipMemcpyAsync(pa, pacc, sizeof(pa), hipMemcpyDeviceToHost, H.stream)) != hipSuccess ||
                (e = hipMemcpyAsync(ph, phist, sizeof(ph), hipMemcpyDeviceToHost, H.stream)) != hipSuccess ||
                (e = hipStreamSynchronize(H.stream)) != hipSuccess)
                return H.hip_fail(e, "precision audit copy");
            float dm, im;
            std::memcpy(&dm, &pa[2], 4);
            std::memcpy(&im, &pa[3], 4);
            stats->precision_class_mismatch += pa[0];
            stats->precision_audited += pa[1];
            stats->precision_dir_err_max = std::max(stats->precision_dir_err_max, (double)dm);
            stats->precision_int_err_max = std::max(stats->precision_int_err_max, (double)im);
            // medians from the 1/4-decade histograms (bin centres)
            auto median = [](const unsigned int* hb) {
                uint64_t tot = 0, run = 0;
                for (int b = 0; b < 64; ++b) tot += hb[b];
                if (!tot) return 0.0;
                int b = 0;
                for (; b < 64; ++b) {
                    run += hb[b];
                    if (2 * run >= tot) break;
                }
                return std::pow(10.0, (b + 0.5) / 4.0 - 12.0);
            };
            stats->precision_dir_err_median = median(ph);
            stats->precision_int_err_median = median(ph + 64);
            if (time_it) {
                float msp = 0;
                (void)hipEventElapsedTime(&msp, H.ev_p0, H.ev_p1);
                stats->precision_audit_ms += msp;
            }
        }
        if (o.use_surrogate) {
            // the network runs on the candidates only
            stats->rays_surrogate += h[5];
            stats->rays_candidates += wk[19];
            stats->rays_gated += wk[20];
            stats->rays_fallback += (uint64_t)wk[19] - std::min<uint64_t>(wk[19], h[5]);
            stats->surrogate_flops += (double)wk[19] * 398848.0;
        }
        if (time_it) {
            float ms = 0, ms2 = 0, ms3 = 0;
            (void)hipEventElapsedTime(&ms, H.ev0, H.ev1);
            (void)hipEventElapsedTime(&ms2, H.ev1, H.ev4);
            if (H.last_band) (void)hipEventElapsedTime(&ms3, H.ev_band0, H.ev_band1);
            stats->kernel_ms += ms;
            stats->escalate_ms += std::max(ms2, ms3);
            if (o.use_surrogate) {
                float ms4 = 0, ms5 = 0;
                (void)hipEventElapsedTime(&ms4, H.ev_s0, H.ev_s1);
                stats->surrogate_ms += ms4;
                if (H.last_audit) {
                    (void)hipEventElapsedTime(&ms5, H.ev_a0, H.ev_a1);
                    stats->audit_ms += ms5;
                }
            }
        }
    }
    return SIRIUS_HIP_OK;
}

}  // namespace sirius

// ---------------------------------------------------------------------------------------
// extern "C" entry points
// ---------------------------------------------------------------------------------------
#define GUARD_BEGIN try {
#define GUARD_END(h)                                                                     \
    }                                                                                    \
    catch (const std::bad_alloc&) { return (h)->fail(SIRIUS_HIP_E_NOMEM, "host allocation"); } \
    catch (...) { return (h)->fail(SIRIUS_HIP_E_DEVICE, "unexpected exception"); }

extern "C" {

int sirius_hip_abi_version(void) { return SIRIUS_HIP_ABI_VERSION; }

SiriusHipHandle sirius_hip_create(void) {
    try {
        return new Handle();
    } catch (...) {
        return nullptr;
    }
}

void sirius_hip_destroy(SiriusHipHandle h) {
    if (!h) return;
    try {
        delete static_cast<Handle*>(h);
    } catch (...) {
    }
}

const char* sirius_hip_last_error(SiriusHipHandle h) {
    if (!h) return "null handle";
    return static_cast<Handle*>(h)->last_error.c_str();
}

int sirius_hip_initialize(SiriusHipHandle hh, int width, int height) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (width <= 0 || height <= 0) return H->fail(SIRIUS_HIP_E_INVALID, "framebuffer size");
    int rc = H->ensure_device();
    if (rc) return rc;
    return sirius_hip_resize(hh, width, height);
    GUARD_END(H)
}

int sirius_hip_resize(SiriusHipHandle hh, int width, int height) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (width <= 0 || height <= 0) return H->fail(SIRIUS_HIP_E_INVALID, "framebuffer size");
    int rc = H->ensure_device();
    if (rc) return rc;
    if (H->fb_dev) (void)hipFree(H->fb_dev);
    H->fb_dev = nullptr;
    size_t bytes = (size_t)width * height * 4 * sizeof(float);
    hipError_t e = hipMalloc(&H->fb_dev, bytes);
    if (e != hipSuccess) { H->fb_dev = nullptr; return H->hip_fail(e, "framebuffer"); }
    // the host copy is pinned: launch's copy of the frame runs at the link's full rate
    if (H->fb_host) (void)hipHostFree(H->fb_host);
    H->fb_host = nullptr;
    H->fb_n = 0;
    const size_t n = (size_t)width * height * 4;
    e = hipHostMalloc((void**)&H->fb_host, n * sizeof(float), hipHostMallocDefault);
    if (e != hipSuccess) { H->fb_host = nullptr; return H->hip_fail(e, "framebuffer host copy"); }
    std::memset(H->fb_host, 0, n * sizeof(float));
    H->fb_n = n;
    H->fb_w = width;
    H->fb_h = height;
    H->initialised = true;
    return SIRIUS_HIP_OK;
    GUARD_END(H)
}

void sirius_hip_cleanup(SiriusHipHandle hh) {
    if (!hh) return;
    Handle* H = static_cast<Handle*>(hh);
    try {
        H->release();
        H->initialised = false;
    } catch (...) {
    }
}

int sirius_hip_is_initialized(SiriusHipHandle hh) {
    return hh && static_cast<Handle*>(hh)->initialised ? 1 : 0;
}

int sirius_hip_device_info(SiriusHipHandle hh, char* name, int name_len, size_t* total_mem, size_t* free_mem,
                           int* cu_count, int* max_threads_per_block) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    int rc = H->ensure_device();
    if (rc) return rc;
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", H->device_name.c_str());
    }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    if (total_mem) *total_mem = tot;
    if (free_mem) *free_mem = fr;
    if (cu_count) *cu_count = H->cu_count;
    if (max_threads_per_block) *max_threads_per_block = H->max_threads;
    return SIRIUS_HIP_OK;
    GUARD_END(H)
}

void sirius_hip_default_opts(SiriusHipTraceOpts* o, int precision) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->integrator = SIRIUS_INTEGRATOR_DP45;
    o->precision = precision;
    o->emission = SIRIUS_EMISSION_THIN_DISK;
    o->schedule = precision == SIRIUS_PRECISION_FP64 ? SIRIUS_SCHEDULE_GRID : SIRIUS_SCHEDULE_PERSISTENT;
    o->rtol = 1e-8;
    o->atol = 1e-10;
    o->h_init = 0.01;
    o->h_min = 1e-8;
    o->h_max = 2.0;
    o->far_slope = 0.0;
    o->symplectic_h = 0.01;
    o->symplectic_order = 4;
    o->max_steps = 20000;
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

int sirius_hip_trace_batch(SiriusHipHandle hh, const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics,
                           int n_metric, const SiriusHipTraceOpts* opts, SiriusHipRayOutput* out,
                           SiriusHipStats* stats) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (!opts || !out) return H->fail(SIRIUS_HIP_E_INVALID, "opts and out are required");
    int rc = H->ensure_device();
    if (rc) return rc;
    Batch B;
    if ((rc = prepare_batch(*H, cams, n_cam, metrics, n_metric, *opts, B))) return rc;
    DevOut dout;
    if ((rc = bind_outputs(*H, out, B.n, opts->n_bins, dout, opts->emission == SIRIUS_EMISSION_VOLUMETRIC))) return rc;
    if (!H->last) H->last = new DevOut;
    *H->last = dout;
    if (out->on_device) {
        // the caller's device arrays are not held: the caller may free them before a later
        // fetch_results / write_rays, which then skip those fields
        DevOut& L = *H->last;
        if (out->cls) L.cls = nullptr;
        if (out->flags) L.flags = nullptr;
        if (out->theta) L.theta = nullptr;
        if (out->phi) L.phi = nullptr;
        if (out->lambda) L.lam = nullptr;
        if (out->disk_r) L.disk_r = nullptr;
        if (out->redshift) L.redshift = nullptr;
        if (out->null_drift) L.null_drift = nullptr;
        if (out->steps) L.steps = nullptr;
        if (out->intensity) L.intensity = nullptr;
        if (out->state) L.state = nullptr;
    }
    H->last_bins = opts->n_bins;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    (void)hipEventRecord(H->ev2, H->stream);
    rc = run_trace(*H, B, *opts, dout, nullptr, nullptr, 0, stats, true);
    if (rc) return rc;
    (void)hipEventRecord(H->ev3, H->stream);
    hipError_t e = hipEventSynchronize(H->ev3);
    if (e != hipSuccess) return H->hip_fail(e, "trace");
    if (stats) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, H->ev2, H->ev3);
        stats->total_ms = ms;
    }
    return copy_outputs_to_host(*H, out, dout, opts->n_bins);
    GUARD_END(H)
}

int sirius_hip_fetch_results(SiriusHipHandle hh, SiriusHipRayOutput* host_out) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (!host_out) return H->fail(SIRIUS_HIP_E_INVALID, "host_out");
    if (!H->last) return H->fail(SIRIUS_HIP_E_STATE, "no trace_batch results to fetch");
    SiriusHipRayOutput o = *host_out;
    o.on_device = 0;
    return copy_outputs_to_host(*H, &o, *H->last, H->last_bins);
    GUARD_END(H)
}

int sirius_hip_last_result_size(SiriusHipHandle hh, int64_t* n_rays, int32_t* n_bins) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    if (!n_rays || !n_bins) return H->fail(SIRIUS_HIP_E_INVALID, "n_rays and n_bins are required");
    if (!H->last) return H->fail(SIRIUS_HIP_E_STATE, "no trace results held");
    *n_rays = H->last->n_rays;
    *n_bins = H->last_bins;
    return SIRIUS_HIP_OK;
}

int sirius_hip_synchronize(SiriusHipHandle hh) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (!H->ready) return SIRIUS_HIP_OK;
    hipError_t e = hipStreamSynchronize(H->stream);
    if (e != hipSuccess) return H->hip_fail(e, "synchronize");
    return SIRIUS_HIP_OK;
    GUARD_END(H)
}

}  // extern "C"
