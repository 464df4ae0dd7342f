// sirius_internal.hpp — host-side state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sirius_hip.h"
#include "geodesic.hpp"

namespace sirius {

// The gates queue rays within this distance (M) of the critical curve, the longest, first, so
// that they all start in the persistent grid's first fill. The 1080p Kerr a = 0.9 frame's tracer
// by the band's width (A/B on one box, ms per frame): none 1.87, 0.25 M 1.34, 0.5 M 1.28,
// 1 M 1.31, 2 M 1.56, 4 M 1.66. A band wider than one fill (5.8 % of that frame lies within
// 0.5 M) queues moderately long rays ahead of the longest.
constexpr double SIRIUS_LONG_BAND = 0.5;

struct DevOut;

struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct MlpState;

// Process-wide memo of host-side constants that depend only on a few doubles (the disk's flux
// maximum, a camera's critical-curve table): keyed by the exact bits of those doubles, bounded
// (cleared when it reaches max_entries), safe to share between handles and threads.
template <int K, class V>
class ConstCache {
  public:
    template <class Make>
    V get(const double (&key)[K], Make make) {
        std::array<uint64_t, K> k;
        std::memcpy(k.data(), key, sizeof(k));
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = map_.find(k);
            if (it != map_.end()) return it->second;
        }
        V v = make();
        std::lock_guard<std::mutex> g(mu_);
        if (map_.size() >= max_entries) map_.clear();
        map_.emplace(k, v);
        return v;
    }
    static constexpr size_t max_entries = 4096;

  private:
    std::mutex mu_;
    std::map<std::array<uint64_t, K>, V> map_;
};

// One handle = one device (device 0), one stream, cached device scratch.
struct Handle {
    bool ready = false;
    bool initialised = false;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // second stream: the FP64 band runs beside the FP32 bulk
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr, ev4 = nullptr;
    hipEvent_t ev_gate = nullptr, ev_band0 = nullptr, ev_band1 = nullptr;
    hipEvent_t ev_s0 = nullptr, ev_s1 = nullptr, ev_a0 = nullptr, ev_a1 = nullptr;  // surrogate, audit
    hipEvent_t ev_p0 = nullptr, ev_p1 = nullptr;  // precision audit
    int cu_count = 0;
    int max_threads = 0;
    size_t total_mem = 0;
    std::string device_name;
    std::string last_error;
    std::map<std::string, Buf> bufs;
    std::vector<float> crit_host;       // host copies that async uploads read from
    std::vector<CamConst> cam_host;
    // device buffers that already hold crit_host / cam_host (an identical batch skips the upload)
    const void* crit_dev_held = nullptr;
    const void* cam_dev_held = nullptr;
    // framebuffer (IAccelerator::initialise / getFrameBuffer)
    float* fb_dev = nullptr;
    float* fb_host = nullptr;  // pinned host copy (getFrameBuffer), fb_n floats
    size_t fb_n = 0;
    int fb_w = 0, fb_h = 0;
    uint8_t* bg_dev = nullptr;
    int bg_w = 0, bg_h = 0;
    // results of the last trace_batch held in device memory (sirius_hip_fetch_results)
    DevOut* last = nullptr;
    int last_bins = 0;
    bool last_band = false;  // the last FP32 trace ran the FP64 band on the second stream
    bool last_audit = false; // the last FP32 trace audited surrogate answers
    // celestial spheres (sirius_hip_set_starfield): device star records, count, brightness
    float* star_dev[2] = {nullptr, nullptr};
    int star_n[2] = {0, 0};
    float star_brightness[2] = {1.f, 1.f};
    // ray bundle of launch (sirius_hip_set_ray_bundle)
    SiriusHipRayBundle bundle = {0, 0, 0.28867513f, 1e-6f, 100.0f, 0};
    // framebuffer shading of launch (sirius_hip_set_shading)
    SiriusHipShading shading = {1, 1, 50000.0, 5.0, 0.01};
    // image writers: the last encoded file and the pinned host staging of converted payloads
    std::vector<unsigned char> image;
    unsigned char* pinned = nullptr;
    size_t pinned_bytes = 0;
    // surrogate network
    MlpState* mlp = nullptr;
    // No-allocation mode (sirius_hip_trace_batch_noalloc): while `arena` is set, scratch()
    // carves named sub-buffers from that fixed device region instead of allocating.
    struct ArenaSlot {
        const char* name;
        size_t off, bytes;
    };
    static constexpr int kArenaSlots = 64;
    unsigned char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    ArenaSlot arena_slots[kArenaSlots];
    int arena_n = 0;

    ~Handle();
    void release();
    int fail(int rc, const std::string& msg);
    int hip_fail(hipError_t e, const char* what);
    int ensure_device();
    void* scratch(const char* name, size_t bytes, int* rc);
};

struct Batch {
    std::vector<CamConst> cc;
    CamConst* cams_dev = nullptr;
    TraceParams tp;
    int family = 0;
    int64_t n = 0;
};

double isco(double a_star);
double photon_orbit(double a, bool prograde);
bool make_cam_const(const SiriusHipCamera& cam, const SiriusHipMetric& m, int64_t offset, CamConst& c,
                    std::string& err);
int prepare_batch(Handle& H, const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics, int n_metric,
                  const SiriusHipTraceOpts& o, Batch& B);
int bind_outputs(Handle& H, const SiriusHipRayOutput* out, int64_t n, int n_bins, DevOut& o, bool volumetric);
int copy_outputs_to_host(Handle& H, const SiriusHipRayOutput* out, const DevOut& o, int n_bins);
int run_trace(Handle& H, Batch& B, const SiriusHipTraceOpts& o, DevOut& dout, const int32_t* list_dev,
              const uint32_t* list_count_dev, int64_t list_cap, SiriusHipStats* stats, bool time_it);
double flops_per_step(int family);

float* upload_critical_tables(Handle& H, Batch& B, int* rc);

// surrogate (mlp_kernels.hip / surrogate.cpp)
void mlp_release(Handle& H);
// Network stage of the FP32 path: gate with the analytic candidate test, fused inference on the
// candidates, per-candidate decision. Fills list_a (FP32 integrator), list_b (FP64 band), list_l
// (taken first) and list_audit; their lengths are work[16], work[17], work[18] and work[10].
int surrogate_stage(Handle& H, Batch& B, const SiriusHipTraceOpts& o, DevOut& dout, int32_t* list_a,
                    int32_t* list_b, int32_t* list_l, int32_t* list_audit, uint32_t* work);
int mlp_forward_dev(Handle& H, const float* x, float* y, int64_t n);

}  // namespace sirius
