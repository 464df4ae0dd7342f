// render_api.cpp — the IAccelerator-facing half of the C-ABI: launch + shade into the RGBA
// framebuffer, background upload, framebuffer access (RDOP006A.h:22-33 roles), the
// critical-curve distance query and the launch-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sirius_hip.h"
#include "geodesic.hpp"
#include "sirius_bundle.hpp"
#include "sirius_internal.hpp"

namespace sirius {
int bundle_for_launch(Handle& H, Batch& B, int emission, int n_cam, BundleOut& bo);
hipError_t launch_star_shade(const DevOut& o, int64_t n, const float* rec0, int n0, float b0, const float* rec1, int n1,
                             float b1, const float* bjac, const uint8_t* bcls, const uint8_t* bflags, float bsize,
                             float* fb, hipStream_t st);
hipError_t launch_shade(const CamConst* cam, const DevOut& o, int64_t n, int emission, const ShadeParams& sp,
                        const uint8_t* bg, int bgw, int bgh, float* fb, hipStream_t st);
hipError_t launch_pixel_average(const float* sub, int w, int h, int gs, float* fb, hipStream_t st);
hipError_t launch_critical_distance(const CamConst* cams, int n_cam, const float* tabs, int64_t n, float* out,
                                    hipStream_t st);
int persistent_blocks_per_cu(bool fp64, int family, int emission, bool multi, int block);

// Radial table of the critical curve (same construction as the oracle's crit_table): spherical
// photon orbits of Kerr-Newman, D = r^2 - 2Mr + a^2 + Q^2, xi(r) = ((r^2 + a^2) D' - 4 r D) /
// (a D'), eta = (r^2 + a^2 - a xi)^2 / D - (xi - a)^2 (Bardeen 1973 with charge), seen from the
// polar angle of the rays' start; out[0] = alpha0, out[1 + j] = rho_c(pi j / (N - 1)).
void critical_table(const CamConst& c, float* out) {
    const int N = SIRIUS_CC_TABLE, S = 4096;
    double M = c.M, a = c.a, Q2 = c.Q2;
    if (std::fabs(a) < 1e-6) {
        double rph = 0.5 * (3 * M + std::sqrt(std::max(9 * M * M - 8 * Q2, 0.0)));
        double D = rph * rph - 2 * M * rph + Q2;
        out[0] = 0.f;
        for (int j = 0; j < N; ++j) out[1 + j] = (float)(rph * rph / std::sqrt(D));
        return;
    }
    double rp = M + std::sqrt(std::max(M * M - a * a - Q2, 0.0));
    double x = c.r_obs * c.er[0], y = c.r_obs * c.er[1], z = c.r_obs * c.er[2];
    double R2 = x * x + y * y + z * z, w = R2 - a * a;
    double r0 = std::sqrt(0.5 * (w + std::sqrt(std::max(w * w + 4 * a * a * z * z, 1e-20))));
    double ct = z / r0, st = std::sqrt(std::max(1 - ct * ct, 1e-300));
    std::vector<double> al, be;
    double amin = 1e300, amax = -1e300;
    for (int i = 0; i < S; ++i) {
        double r = rp + (5.0 * M - rp) * (i + 0.5) / S;
        double D = r * r - 2 * M * r + a * a + Q2, Dp = 2 * r - 2 * M;
        if (D <= 0 || std::fabs(Dp) < 1e-12) continue;
        double xi = ((r * r + a * a) * Dp - 4 * r * D) / (a * Dp);
        double t1 = r * r + a * a - a * xi;
        double eta = t1 * t1 / D - (xi - a) * (xi - a);
        if (eta < 0) continue;
        double b2 = eta + a * a * ct * ct - xi * xi * ct * ct / (st * st);
        if (b2 < 0) continue;
        al.push_back(-xi / st);
        be.push_back(std::sqrt(b2));
        amin = std::min(amin, al.back());
        amax = std::max(amax, al.back());
    }
    double c0 = 0.5 * (amin + amax);
    out[0] = (float)c0;
    std::vector<std::pair<double, double>> pr;
    for (size_t i = 0; i < al.size(); ++i) pr.emplace_back(std::atan2(be[i], al[i] - c0), std::hypot(al[i] - c0, be[i]));
    std::stable_sort(pr.begin(), pr.end(), [](auto& u, auto& v) { return u.first < v.first; });
    size_t k = 0, n = pr.size();
    for (int j = 0; j < N; ++j) {
        double psi = M_PI * j / (N - 1), v = 0;
        if (n == 0) v = 0;
        else if (psi <= pr[0].first) v = pr[0].second;
        else if (psi >= pr[n - 1].first) v = pr[n - 1].second;
        else {
            while (k + 2 < n && pr[k + 1].first < psi) ++k;
            double t = (psi - pr[k].first) / (pr[k + 1].first - pr[k].first);
            v = pr[k].second + t * (pr[k + 1].second - pr[k].second);
        }
        out[1 + j] = (float)v;
    }
}

// Uploads the critical tables of a batch (n_cam x (N + 1) floats) and returns the device copy.
float* upload_critical_tables(Handle& H, Batch& B, int* rc) {
    int n_cam = (int)B.cc.size();
    // the host copy lives in the handle, so the asynchronous upload needs no synchronisation
    // (the stream is synchronised before this buffer is rewritten by the next batch)
    static ConstCache<7, std::vector<float>> table_cache;
    const size_t per = SIRIUS_CC_TABLE + 1;
    std::vector<float> tabs((size_t)n_cam * per, 0.f);
    for (int i = 0; i < n_cam; ++i) {
        const CamConst& c = B.cc[i];
        if (c.family != 0) continue;
        const double key[7] = {c.M, c.a, c.Q2, c.r_obs, c.er[0], c.er[1], c.er[2]};
        std::vector<float> t = table_cache.get(key, [&] {
            std::vector<float> v(per);
            critical_table(c, v.data());
            return v;
        });
        std::memcpy(tabs.data() + (size_t)i * per, t.data(), per * 4);
    }
    float* d = (float*)H.scratch("crit_tabs", tabs.size() * 4, rc);
    if (*rc) return nullptr;
    // batches end synchronised, so the device still holds the last upload: skip an identical one
    if (H.crit_dev_held == d && H.crit_host == tabs) return d;
    H.crit_host.swap(tabs);
    H.crit_dev_held = nullptr;
    hipError_t e = hipMemcpyAsync(d, H.crit_host.data(), H.crit_host.size() * 4, hipMemcpyHostToDevice, H.stream);
    if (e != hipSuccess) {
        *rc = H.hip_fail(e, "upload critical tables");
        return nullptr;
    }
    H.crit_dev_held = d;
    return d;
}

int critical_distance_dev(Handle& H, Batch& B, float* out_dev) {
    int rc = 0;
    float* tabs = upload_critical_tables(H, B, &rc);
    if (rc) return rc;
    hipError_t e = launch_critical_distance(B.cams_dev, (int)B.cc.size(), tabs, B.n, out_dev, H.stream);
    if (e != hipSuccess) return H.hip_fail(e, "critical distance");
    return SIRIUS_HIP_OK;
}

}  // namespace sirius

using namespace sirius;

#define GUARD_BEGIN try {
#define GUARD_END(h)                                                                     \
    }                                                                                    \
    catch (const std::bad_alloc&) { return (h)->fail(SIRIUS_HIP_E_NOMEM, "host allocation"); } \
    catch (...) { return (h)->fail(SIRIUS_HIP_E_DEVICE, "unexpected exception"); }

extern "C" {

void sirius_hip_default_shading(SiriusHipShading* s) {
    if (!s) return;
    s->samples_per_pixel = 1;
    s->temperature_model = 1;
    s->disk_temperature = 50000.0;
    s->disk_boost = 5.0;
    s->max_steps_level = 0.01;
}

int sirius_hip_set_shading(SiriusHipHandle hh, const SiriusHipShading* s) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    if (!s) return H->fail(SIRIUS_HIP_E_INVALID, "shading");
    if (s->samples_per_pixel < 1 || s->samples_per_pixel > 4096)
        return H->fail(SIRIUS_HIP_E_INVALID, "samples_per_pixel must be in [1, 4096]");
    if (s->temperature_model != 0 && s->temperature_model != 1)
        return H->fail(SIRIUS_HIP_E_INVALID, "temperature_model is 0 (Shakura-Sunyaev) or 1 (Novikov-Thorne)");
    if (!(s->disk_temperature >= 0) || !(s->disk_boost >= 0) || !(s->max_steps_level >= 0))
        return H->fail(SIRIUS_HIP_E_INVALID, "disk_temperature, disk_boost and max_steps_level must be >= 0");
    H->shading = *s;
    return SIRIUS_HIP_OK;
}

int sirius_hip_launch(SiriusHipHandle hh, const SiriusHipCamera* cam, const SiriusHipMetric* metric,
                      const SiriusHipTraceOpts* opts) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (!H->initialised) return H->fail(SIRIUS_HIP_E_STATE, "sirius_hip_initialize first");
    if (!cam || !metric || !opts) return H->fail(SIRIUS_HIP_E_INVALID, "camera, metric and opts are required");
    if (cam->width != H->fb_w || cam->height != H->fb_h)
        return H->fail(SIRIUS_HIP_E_INVALID, "camera size differs from the framebuffer (resize first)");
    // g x g rays per pixel (shadePixel, SNRS001A.cpp:484-495): the camera of g W x g H pixels has
    // its pixel (g px + sx, g py + sy) centred on (px + (sx + 1/2) / g, py + (sy + 1/2) / g) of this
    // one, at the same field of view
    const int gs = std::max(1, (int)std::floor(std::sqrt((double)H->shading.samples_per_pixel) + 1e-9));
    SiriusHipCamera c = *cam;
    c.metric_index = 0;
    c.width *= gs;
    c.height *= gs;
    Batch B;
    int rc = prepare_batch(*H, &c, 1, metric, 1, *opts, B);
    if (rc) return rc;
    DevOut dout;
    if ((rc = bind_outputs(*H, nullptr, B.n, opts->n_bins, dout, opts->emission == SIRIUS_EMISSION_VOLUMETRIC))) return rc;
    // the frame's rays overwrite the library's result arrays: they are now the last results
    // (sirius_hip_fetch_results, sirius_hip_write_rays)
    if (!H->last) H->last = new DevOut;
    *H->last = dout;
    H->last_bins = opts->n_bins;
    rc = run_trace(*H, B, *opts, dout, nullptr, nullptr, 0, nullptr, false);
    if (rc) return rc;
    // radiance per ray (into the framebuffer itself at one ray per pixel), star light added per
    // ray, then each pixel's average
    float* sub = H->fb_dev;
    if (gs > 1) {
        sub = (float*)H->scratch("fb_sub", (size_t)B.n * 4 * sizeof(float), &rc);
        if (rc) return rc;
    }
    const ShadeParams sp{H->shading.temperature_model, H->shading.disk_temperature, H->shading.disk_boost,
                         H->shading.max_steps_level};
    hipError_t e = launch_shade(B.cams_dev, dout, B.n, opts->emission, sp, H->bg_dev, H->bg_w, H->bg_h, sub, H->stream);
    if (e != hipSuccess) return H->hip_fail(e, "shade");
    // with a ray bundle, star light is averaged over each pixel's beam (the bundle pass traces
    // the frame's rays again in FP64 with their deviation)
    BundleOut bo{};
    const bool stars = (H->star_dev[0] && H->star_n[0] > 0) || (H->star_dev[1] && H->star_n[1] > 0);
    if (H->bundle.enabled && stars && opts->integrator == SIRIUS_INTEGRATOR_DP45) {
        const int emission = opts->emission == SIRIUS_EMISSION_THIN_DISK ? SIRIUS_EMISSION_THIN_DISK : 0;
        if ((rc = bundle_for_launch(*H, B, emission, 1, bo))) return rc;
    }
    e = launch_star_shade(dout, B.n, H->star_dev[0], H->star_n[0], H->star_brightness[0], H->star_dev[1],
                          H->star_n[1], H->star_brightness[1], bo.jac, bo.cls, bo.flags, H->bundle.initial_size,
                          sub, H->stream);
    if (e != hipSuccess) return H->hip_fail(e, "star shade");
    if (gs > 1 && (e = launch_pixel_average(sub, cam->width, cam->height, gs, H->fb_dev, H->stream)) != hipSuccess)
        return H->hip_fail(e, "pixel average");
    e = hipMemcpyAsync(H->fb_host, H->fb_dev, H->fb_n * sizeof(float), hipMemcpyDeviceToHost,
                       H->stream);
    if (e != hipSuccess) return H->hip_fail(e, "framebuffer copy");
    e = hipStreamSynchronize(H->stream);  // launch is synchronous (RDOP001A.cu:1044)
    if (e != hipSuccess) return H->hip_fail(e, "synchronize");
    return SIRIUS_HIP_OK;
    GUARD_END(H)
}

float* sirius_hip_get_frame_buffer(SiriusHipHandle hh) {
    if (!hh) return nullptr;
    Handle* H = static_cast<Handle*>(hh);
    return H->initialised && H->fb_host ? H->fb_host : nullptr;
}

size_t sirius_hip_get_frame_buffer_size(SiriusHipHandle hh) {
    if (!hh) return 0;
    Handle* H = static_cast<Handle*>(hh);
    return (size_t)H->fb_w * H->fb_h * 4 * sizeof(float);
}

int sirius_hip_upload_background(SiriusHipHandle hh, const unsigned char* rgba, int width, int height) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (!rgba || width <= 0 || height <= 0) return H->fail(SIRIUS_HIP_E_INVALID, "background image");
    int rc = H->ensure_device();
    if (rc) return rc;
    (void)hipStreamSynchronize(H->stream);  // a shading launch may still read the old image
    if (H->bg_dev) (void)hipFree(H->bg_dev);
    H->bg_dev = nullptr;
    size_t bytes = (size_t)width * height * 4;
    hipError_t e = hipMalloc(&H->bg_dev, bytes);
    if (e != hipSuccess) { H->bg_dev = nullptr; return H->hip_fail(e, "background"); }
    e = hipMemcpy(H->bg_dev, rgba, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return H->hip_fail(e, "background upload");
    H->bg_w = width;
    H->bg_h = height;
    return SIRIUS_HIP_OK;
    GUARD_END(H)
}

int sirius_hip_critical_distance(SiriusHipHandle hh, const SiriusHipCamera* cams, int n_cam,
                                 const SiriusHipMetric* metrics, int n_metric, float* out, int on_device) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (!out) return H->fail(SIRIUS_HIP_E_INVALID, "out");
    int rc = H->ensure_device();
    if (rc) return rc;
    SiriusHipTraceOpts o;
    sirius_hip_default_opts(&o, SIRIUS_PRECISION_FP64);
    Batch B;
    if ((rc = prepare_batch(*H, cams, n_cam, metrics, n_metric, o, B))) return rc;
    float* d = on_device ? out : (float*)H->scratch("crit_out", 4 * (size_t)B.n, &rc);
    if (rc) return rc;
    if ((rc = critical_distance_dev(*H, B, d))) return rc;
    if (!on_device) {
        // the kernel ran on the library's non-blocking stream: copy on that stream, then wait
        hipError_t e = hipMemcpyAsync(out, d, 4 * (size_t)B.n, hipMemcpyDeviceToHost, H->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(H->stream);
        if (e != hipSuccess) return H->hip_fail(e, "copy");
    } else {
        hipError_t e = hipStreamSynchronize(H->stream);
        if (e != hipSuccess) return H->hip_fail(e, "synchronize");
    }
    return SIRIUS_HIP_OK;
    GUARD_END(H)
}

// Autotuner: times the FP32 persistent trace over block sizes {64, 128, 256} x refill threshold
// {8, 16, 24, 32}, each at full occupancy (blocks per CU from the occupancy API), and keeps the
// fastest. Batches of up to 4 M rays are timed whole; larger ones on a strided sample (every
// k-th ray, about 1/16 of the batch, at least 64k rays).
int sirius_hip_autotune(SiriusHipHandle hh, const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics,
                        int n_metric, SiriusHipTraceOpts* opts, double* best_ms) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
    if (!opts) return H->fail(SIRIUS_HIP_E_INVALID, "opts");
    int rc = H->ensure_device();
    if (rc) return rc;
    if (opts->precision != SIRIUS_PRECISION_FP32_ESCALATE || opts->integrator != SIRIUS_INTEGRATOR_DP45)
        return H->fail(SIRIUS_HIP_E_UNSUPPORTED, "autotune applies to the FP32 persistent DP45 path");
    Batch B;
    if ((rc = prepare_batch(*H, cams, n_cam, metrics, n_metric, *opts, B))) return rc;
    const bool whole = B.n <= ((int64_t)1 << 22);
    int64_t stride = std::max<int64_t>(1, B.n / std::max<int64_t>(65536, B.n / 16));
    std::vector<int32_t> ids;
    for (int64_t i = 0; i < B.n; i += stride) ids.push_back((int32_t)i);
    int32_t* list = (int32_t*)H->scratch("tune_list", 4 * ids.size(), &rc);
    uint32_t* cnt = (uint32_t*)H->scratch("tune_cnt", 4, &rc);
    if (rc) return rc;
    uint32_t nn = (uint32_t)ids.size();
    hipError_t e;
    if ((e = hipStreamSynchronize(H->stream)) != hipSuccess) return H->hip_fail(e, "synchronize");
    if ((e = hipMemcpy(list, ids.data(), 4 * ids.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return H->hip_fail(e, "upload");
    if ((e = hipMemcpy(cnt, &nn, 4, hipMemcpyHostToDevice)) != hipSuccess) return H->hip_fail(e, "upload");
    DevOut dout;
    if ((rc = bind_outputs(*H, nullptr, B.n, opts->n_bins, dout, opts->emission == SIRIUS_EMISSION_VOLUMETRIC))) return rc;
    double best = 1e300;
    SiriusHipTraceOpts bo = *opts;
    const int blocks[3] = {64, 128, 256};
    const int refills[4] = {8, 16, 24, 32};  // 1 and 4 never won (1080p Kerr frame: 1.3-1.5 ms against 1.1)
    // every block size at full occupancy: on a sample, lower occupancies can time faster (its
    // short trace has little latency to hide) while on whole batches they ran up to 1.7x slower
    for (int bs : blocks) {
        int occ = persistent_blocks_per_cu(false, B.family, B.tp.emission, B.tp.n_cam > 1, bs);
        int bpcs[1] = {occ};
        for (int j = 0; j < 1; ++j) {
            for (int rf : refills) {
                SiriusHipTraceOpts t = *opts;
                t.block_size = bs;
                t.blocks_per_cu = bpcs[j];
                t.refill_min = rf;
                t.use_surrogate = 0;
                // the whole batch when it is up to 4 M rays (a sample's short trace ranks launch
                // shapes differently from a whole frame's), the strided sample beyond; the
                // faster of two timings
                double ms = 1e300;
                for (int rep = 0; rep < 2; ++rep) {
                    SiriusHipStats st;
                    std::memset(&st, 0, sizeof(st));
                    rc = whole ? run_trace(*H, B, t, dout, nullptr, nullptr, 0, &st, true)
                               : run_trace(*H, B, t, dout, list, cnt, (int64_t)ids.size(), &st, true);
                    if (rc) return rc;
                    ms = std::min(ms, st.kernel_ms);
                }
                if (ms < best) {
                    best = ms;
                    bo = t;
                }
            }
        }
    }
    opts->block_size = bo.block_size;
    opts->blocks_per_cu = bo.blocks_per_cu;
    opts->refill_min = bo.refill_min;
    if (best_ms) *best_ms = best;
    return SIRIUS_HIP_OK;
    GUARD_END(H)
}

}  // extern "C"
