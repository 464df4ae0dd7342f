// noalloc.hip — sirius_hip_trace_batch_noalloc: the batched trace of sirius_hip_trace_batch with
// every device scratch buffer (camera constants, critical-curve tables, work block, work lists,
// per-ray final states and step counters, statistics accumulators) carved from one fixed device
// array instead of allocated, for callers that cannot call hipMalloc (code under graph capture,
// tight serving loops, harnesses that forbid allocation). The kernels are the ones trace_batch
// launches, with the same arguments; only where their scratch lives differs.
//
// The array is a file-scope __device__ object: every GPU the library runs on holds its own copy
// until the process ends, and every call on that GPU (from any handle) shares it. Calls that
// overlap on one GPU race on it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/sirius_hip.h"
#include "sirius_internal.hpp"

// 256 MiB: a 1920x1080 frame fits with the caller supplying the result arrays (88 B of scratch per
// ray: the final state, two step counters and four work lists).
static constexpr size_t kNoallocScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
unsigned char* g_scratch_dev[kMaxDevices] = {};  // the array's device address, per device
std::mutex g_scratch_mu;

// Device address of the current device's copy of `scratch`, looked up once per device with that
// device current. The array's name is a host-side stand-in and never reaches a kernel or a copy.
int scratch_address(sirius::Handle& H, unsigned char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return H.hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return H.fail(SIRIUS_HIP_E_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return H.hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return H.fail(SIRIUS_HIP_E_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_dev[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_dev[dev];
    return SIRIUS_HIP_OK;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Upper bound of the scratch a batch takes: the buffers run_trace and its helpers request by name
// (each rounded up to 256 bytes), for the arrays the caller does not supply.
size_t scratch_needed(const sirius::Batch& B, const SiriusHipTraceOpts& o, const SiriusHipRayOutput& out) {
    const size_t n = (size_t)B.n, nc = B.cc.size(), nb = (size_t)(o.n_bins > 0 ? o.n_bins : 1);
    size_t need = up256(sizeof(sirius::CamConst) * nc) + up256(128);  // cameras, work block
    need += out.cls ? 0 : up256(n);
    need += out.flags ? 0 : up256(n);
    for (const void* p : {(const void*)out.theta, (const void*)out.phi, (const void*)out.lambda,
                          (const void*)out.disk_r, (const void*)out.redshift, (const void*)out.null_drift,
                          (const void*)out.steps})
        need += p ? 0 : up256(4 * n);
    need += out.intensity ? 0 : up256(4 * n * nb);
    need += out.state ? 0 : up256(64 * n);
    need += 2 * up256(4 * n);  // rejected and landing counters
    if (o.emission == SIRIUS_EMISSION_VOLUMETRIC && B.family == SIRIUS_METRIC_KERR_SCHILD) need += up256(8 * n * nb);
    const bool fp32_queue = o.precision == SIRIUS_PRECISION_FP32_ESCALATE && o.schedule == SIRIUS_SCHEDULE_PERSISTENT &&
                            o.integrator == SIRIUS_INTEGRATOR_DP45;
    if (fp32_queue) {
        need += up256(4 * n);  // escalation list
        if (B.family == SIRIUS_METRIC_KERR_SCHILD)
            need += 3 * up256(4 * n) + up256(4 * (SIRIUS_CC_TABLE + 1) * nc);  // gate lists, critical tables
    }
    need += up256(64) + up256(4 * sirius::SIRIUS_DRIFT_BINS);  // statistics accumulators
    return need;
}

// Leaves the handle in its ordinary (allocating) mode whichever way the call returns. The uploads
// memoised during the call live in the shared array, which another call may overwrite: forget them.
struct ArenaScope {
    sirius::Handle* H;
    ~ArenaScope() {
        H->arena = nullptr;
        H->arena_cap = H->arena_used = 0;
        H->arena_n = 0;
        H->cam_dev_held = H->crit_dev_held = nullptr;
    }
};

}  // namespace

extern "C" int sirius_hip_trace_batch_noalloc(SiriusHipHandle hh, const SiriusHipCamera* cams, int n_cam,
                                              const SiriusHipMetric* metrics, int n_metric,
                                              const SiriusHipTraceOpts* opts, SiriusHipRayOutput* out,
                                              SiriusHipStats* stats) {
    using namespace sirius;
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    try {
        if (!opts || !out) return H->fail(SIRIUS_HIP_E_INVALID, "opts and out are required");
        if (!out->on_device)
            return H->fail(SIRIUS_HIP_E_INVALID, "the no-allocation trace writes device arrays only (on_device = 1)");
        if (opts->use_surrogate)
            return H->fail(SIRIUS_HIP_E_UNSUPPORTED, "the no-allocation trace does not run the surrogate");
        int rc = H->ensure_device();
        if (rc) return rc;
        unsigned char* base = nullptr;
        if ((rc = scratch_address(*H, &base))) return rc;
        // the results of this call are never held: they live in the caller's arrays or in the
        // shared scratch, which the next call on this GPU overwrites
        delete H->last;
        H->last = nullptr;
        H->arena = base;
        H->arena_cap = kNoallocScratchBytes;
        H->arena_used = 0;
        H->arena_n = 0;
        H->cam_dev_held = H->crit_dev_held = nullptr;
        ArenaScope scope{H};
        Batch B;
        if ((rc = prepare_batch(*H, cams, n_cam, metrics, n_metric, *opts, B))) return rc;
        const size_t need = scratch_needed(B, *opts, *out);
        if (need > kNoallocScratchBytes)
            return H->fail(SIRIUS_HIP_E_INVALID, "batch of " + std::to_string(B.n) + " rays needs " +
                                                     std::to_string((need + (1 << 20) - 1) >> 20) +
                                                     " MiB of scratch; the no-allocation scratch holds " +
                                                     std::to_string(kNoallocScratchBytes >> 20) + " MiB");
        DevOut dout;
        if ((rc = bind_outputs(*H, out, B.n, opts->n_bins, dout, opts->emission == SIRIUS_EMISSION_VOLUMETRIC)))
            return rc;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        (void)hipEventRecord(H->ev2, H->stream);
        if ((rc = run_trace(*H, B, *opts, dout, nullptr, nullptr, 0, stats, stats != nullptr))) return rc;
        (void)hipEventRecord(H->ev3, H->stream);
        hipError_t e = hipEventSynchronize(H->ev3);
        if (e != hipSuccess) return H->hip_fail(e, "trace");
        if (stats) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, H->ev2, H->ev3);
            stats->total_ms = ms;
        }
        return SIRIUS_HIP_OK;
    } catch (const std::bad_alloc&) {
        return H->fail(SIRIUS_HIP_E_NOMEM, "host allocation");
    } catch (...) {
        return H->fail(SIRIUS_HIP_E_DEVICE, "unexpected exception");
    }
}
