// bundle_api.cpp — host side of the ray-bundle pass (sirius_hip_trace_bundle,
// sirius_hip_set_ray_bundle) and its use by sirius_hip_launch (bundle_for_launch).
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/sirius_hip.h"
#include "sirius_bundle.hpp"
#include "sirius_internal.hpp"

namespace sirius {

namespace {

int check_bundle(Handle& H, const SiriusHipRayBundle& b) {
    if (!(b.initial_size > 0.0f && b.initial_size < 1e3f) || !(b.min_bundle_size >= 0.0f) ||
        !(b.max_ellipse_ratio >= 1.0f))
        return H.fail(SIRIUS_HIP_E_INVALID, "ray bundle: initial_size in (0, 1000), min_bundle_size >= 0, "
                                            "max_ellipse_ratio >= 1");
    return SIRIUS_HIP_OK;
}

BundleParams params_of(const SiriusHipRayBundle& b) {
    BundleParams p;
    p.size = b.initial_size;
    p.min_size = b.min_bundle_size;
    p.max_ratio = b.max_ellipse_ratio;
    return p;
}

}  // namespace

// The bundle pass over a prepared batch into library scratch (launch's beams).
int bundle_for_launch(Handle& H, Batch& B, int emission, int n_cam, BundleOut& bo) {
    int rc = SIRIUS_HIP_OK;
    bo.n = B.n;
    bo.cls = (uint8_t*)H.scratch("b_cls", B.n, &rc);
    if (!rc) bo.flags = (uint8_t*)H.scratch("b_flags", B.n, &rc);
    if (!rc) bo.jac = (float*)H.scratch("b_jac", 16 * (size_t)B.n, &rc);
    if (!rc) bo.beam = (float*)H.scratch("b_beam", 12 * (size_t)B.n, &rc);
    if (rc) return rc;
    hipError_t e = launch_trace_bundle(B.family, emission, B.cams_dev, n_cam, B.tp, params_of(H.bundle), bo, H.stream);
    if (e != hipSuccess) return H.hip_fail(e, "bundle trace");
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

void sirius_hip_default_ray_bundle(SiriusHipRayBundle* b) {
    if (!b) return;
    std::memset(b, 0, sizeof(*b));
    b->enabled = 0;
    b->use_full_riemann = 0;
    b->initial_size = 0.28867513f;
    b->min_bundle_size = 1e-6f;
    b->max_ellipse_ratio = 100.0f;
}

int sirius_hip_set_ray_bundle(SiriusHipHandle hh, const SiriusHipRayBundle* b) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    if (!b) return H->fail(SIRIUS_HIP_E_INVALID, "ray bundle parameters");
    int rc = check_bundle(*H, *b);
    if (rc) return rc;
    H->bundle = *b;
    return SIRIUS_HIP_OK;
}

int sirius_hip_trace_bundle(SiriusHipHandle hh, const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics,
                            int n_metric, const SiriusHipTraceOpts* opts, SiriusHipBundleOutput* out) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
        if (!opts || !out) return H->fail(SIRIUS_HIP_E_INVALID, "opts and out are required");
        if (opts->integrator != SIRIUS_INTEGRATOR_DP45)
            return H->fail(SIRIUS_HIP_E_UNSUPPORTED, "ray bundles follow the DP45 integrator");
        int rc = H->ensure_device();
        if (rc) return rc;
        SiriusHipTraceOpts o = *opts;
        o.use_surrogate = 0;
        o.precision = SIRIUS_PRECISION_FP64;
        Batch B;
        if ((rc = prepare_batch(*H, cams, n_cam, metrics, n_metric, o, B))) return rc;
        const int64_t n = B.n;
        BundleOut bo;
        bo.n = n;
        auto get = [&](const char* name, void* user, size_t bytes) -> void* {
            if (out->on_device && user) return user;
            return H->scratch(name, bytes, &rc);
        };
        bo.cls = (uint8_t*)get("b_cls", out->cls, n);
        if (!rc) bo.flags = (uint8_t*)get("b_flags", out->flags, n);
        if (!rc) bo.jac = (float*)get("b_jac", out->jac, 16 * (size_t)n);
        if (!rc) bo.beam = (float*)get("b_beam", out->beam, 12 * (size_t)n);
        if (rc) return rc;
        const int emission = opts->emission == SIRIUS_EMISSION_THIN_DISK ? SIRIUS_EMISSION_THIN_DISK : 0;
        hipError_t e = launch_trace_bundle(B.family, emission, B.cams_dev, n_cam, B.tp, params_of(H->bundle), bo,
                                           H->stream);
        if (e != hipSuccess) return H->hip_fail(e, "bundle trace");
        if (!out->on_device) {
            struct { void* dst; const void* src; size_t bytes; } c[] = {
                {out->cls, bo.cls, (size_t)n}, {out->flags, bo.flags, (size_t)n},
                {out->jac, bo.jac, 16 * (size_t)n}, {out->beam, bo.beam, 12 * (size_t)n}};
            for (auto& x : c)
                if (x.dst && (e = hipMemcpyAsync(x.dst, x.src, x.bytes, hipMemcpyDeviceToHost, H->stream)) != hipSuccess)
                    return H->hip_fail(e, "bundle results copy");
        }
        if ((e = hipStreamSynchronize(H->stream)) != hipSuccess) return H->hip_fail(e, "bundle trace");
        return SIRIUS_HIP_OK;
    GUARD_END(H)
}

}  // extern "C"
