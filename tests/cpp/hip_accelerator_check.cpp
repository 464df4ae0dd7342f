// hip_accelerator_check.cpp — drives HIPAccelerator (include/sirius_hip_accelerator.hpp) the way a
// Sirius tree does: initialise, launch with a LaunchConfig, read the framebuffer, the star field
// of LaunchConfig::enableStarfield with ray bundles, and the image writers. Built by
// __graft_entry__.build(); run by tests/test_gpu_cpp_wrapper.py. Prints one line per check and
// exits nonzero at the first failure.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>


#include "sirius_hip_accelerator.hpp"

using Sirius::Acceleration::LaunchConfig;
using Sirius::Acceleration::HIP::HIPAccelerator;

static int fail(const char* what, const HIPAccelerator& a) {
    std::printf("FAIL %s: %s\n", what, a.getLastError().c_str());
    return 1;
}

static double escaped_light(const float* fb, const float* base, size_t n) {
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += (double)fb[i] - (double)base[i];
    return s;
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    HIPAccelerator acc;
    const int w = 96, h = 54;
    if (!acc.initialise(w, h)) return fail("initialise", acc);
    LaunchConfig cfg;
    cfg.width = w;
    cfg.height = h;
    cfg.blackHoleMass = 1.0f;
    cfg.blackHoleSpin = 0.9f;
    cfg.observerDistance = 50.0f;
    cfg.observerInclination = 1.309f;  // 75 degrees
    cfg.cameraFOV = 20.0f;
    cfg.diskInnerRadius = 0.0f;
    cfg.diskOuterRadius = 30.0f;
    cfg.metricFamily = 0;
    acc.launch(cfg);
    const float* fb = acc.getFrameBuffer();
    if (!fb || acc.getFrameBufferSize() != (size_t)w * h * 16) return fail("framebuffer", acc);
    float mx = 0;
    bool finite = true;
    for (int i = 0; i < w * h * 4; ++i) {
        if (i % 4 != 3) mx = std::fmax(mx, fb[i]);  // colour channels (alpha is 1 everywhere)
        finite &= std::isfinite(fb[i]);
    }
    if (!(mx > 0) || !finite) return fail("disk frame lit and finite", acc);
    std::printf("ok launch: max %.4g\n", mx);

    // metricType names the spacetime (RDOP003A.h:63-75): Schwarzschild ignores the spin
    cfg.samplesPerPixel = 1;
    cfg.metricType = 1;
    acc.launch(cfg);
    std::vector<float> schw(acc.getFrameBuffer(), acc.getFrameBuffer() + (size_t)w * h * 4);
    cfg.metricType = 2;
    cfg.blackHoleSpin = 0.0f;
    acc.launch(cfg);
    if (std::vector<float>(acc.getFrameBuffer(), acc.getFrameBuffer() + (size_t)w * h * 4) != schw)
        return fail("metricType 1 is Kerr with a = 0", acc);
    cfg.blackHoleSpin = 0.9f;
    acc.launch(cfg);
    if (std::vector<float>(acc.getFrameBuffer(), acc.getFrameBuffer() + (size_t)w * h * 4) == schw)
        return fail("metricType 2 keeps the spin", acc);
    cfg.metricType = 4;  // Goedel: not a spacetime of this backend
    acc.launch(cfg);
    if (acc.getLastError().find("metricType 4") == std::string::npos) return fail("metricType 4 refused", acc);
    cfg.metricType = 7;  // Ellis drainhole with a Kerr-Schild family: inconsistent
    acc.launch(cfg);
    if (acc.getLastError().find("metricFamily") == std::string::npos) return fail("metricType 7 / family 0", acc);
    cfg.metricType = 2;
    std::printf("ok metricType\n");

    // samplesPerPixel: a 2 x 2 grid per pixel averages four rays; the frame changes, its mean
    // brightness barely
    acc.launch(cfg);
    std::vector<float> one(acc.getFrameBuffer(), acc.getFrameBuffer() + (size_t)w * h * 4);
    cfg.samplesPerPixel = 4;
    acc.launch(cfg);
    std::vector<float> four(acc.getFrameBuffer(), acc.getFrameBuffer() + (size_t)w * h * 4);
    double s1 = 0, s4 = 0;
    for (size_t i = 0; i < one.size(); ++i)
        if (i % 4 != 3) { s1 += one[i]; s4 += four[i]; }
    if (one == four || !(std::fabs(s4 - s1) < 0.2 * s1)) return fail("samplesPerPixel", acc);
    std::printf("ok samplesPerPixel: mean %.4g at 1, %.4g at 4\n", s1 / (w * h), s4 / (w * h));
    cfg.samplesPerPixel = 1;

    // a wide view of the sky: the default catalogue's stars, averaged over each pixel's beam
    cfg.cameraFOV = 100.0f;
    cfg.diskOuterRadius = 0.5f;  // no disk in the way
    acc.launch(cfg);
    std::vector<float> base(acc.getFrameBuffer(), acc.getFrameBuffer() + (size_t)w * h * 4);
    cfg.enableStarfield = true;
    cfg.starfieldBrightness = 1.0f;
    if (!acc.setRayBundle(true)) return fail("setRayBundle", acc);
    acc.launch(cfg);
    const double lit = escaped_light(acc.getFrameBuffer(), base.data(), base.size());
    if (!(lit > 0)) return fail("star light with the bundle", acc);
    cfg.starfieldBrightness = 2.0f;  // regenerated with the new brightness: twice the light
    acc.launch(cfg);
    const double lit2 = escaped_light(acc.getFrameBuffer(), base.data(), base.size());
    if (!(std::fabs(lit2 - 2 * lit) <= 1e-4 * lit2)) return fail("brightness scales the light", acc);
    cfg.enableStarfield = false;
    acc.launch(cfg);
    if (escaped_light(acc.getFrameBuffer(), base.data(), base.size()) != 0.0) return fail("stars removed", acc);
    std::printf("ok starfield: light %.6g, at brightness 2: %.6g\n", lit, lit2);

    for (int fmt : {SIRIUS_IMAGE_EXR, SIRIUS_IMAGE_PNG, SIRIUS_IMAGE_PPM}) {
        const std::string path = dir + "/frame." + std::to_string(fmt);
        if (!acc.writeImage(path, fmt, 1.0f)) return fail("writeImage", acc);
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) return fail("image file", acc);
        std::fseek(f, 0, SEEK_END);
        const long sz = std::ftell(f);
        std::fclose(f);
        if (sz <= 0) return fail("image size", acc);
        std::printf("ok image %d: %ld bytes\n", fmt, sz);
    }
    acc.cleanup();
    std::printf("OK\n");
    return 0;
}
