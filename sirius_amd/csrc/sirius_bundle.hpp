// sirius_bundle.hpp — ray-bundle pass: parameters and result arrays shared by the bundle
// kernel (bundle_kernels.hip), the star shading (util_kernels.hip) and the host code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/sirius_hip.h"
#include "geodesic.hpp"

namespace sirius {

struct BundleParams {
    double size;      // beam radius at the camera, pixels
    float min_size;   // beams with a smaller semi-major axis are terminated
    float max_ratio;  // beams more elongated than this are terminated
};

// Per-ray results (structure of arrays, n entries; jac 4 n and beam 3 n, component-major).
struct BundleOut {
    uint8_t* cls;
    uint8_t* flags;
    float* jac;   // d(a, b)/d(u, v): da/du, da/dv, db/du, db/dv
    float* beam;  // semi-major, semi-minor, orientation
    int64_t n;
};

hipError_t launch_trace_bundle(int family, int emission, const CamConst* cams, int n_cam, const TraceParams& tp,
                               const BundleParams& bp, const BundleOut& bo, hipStream_t st);

}  // namespace sirius
