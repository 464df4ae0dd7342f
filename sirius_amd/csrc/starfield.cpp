// starfield.cpp — procedural celestial spheres.
//
// The catalogue generator restates StarfieldGenerator::generate / generateStar
// (src/Sirius.Core/Environment/PHSF001A.h:133-151, 206-249) on the host: the same draws from
// std::mt19937 through std::uniform_real_distribution<float>, in the same order and with the
// same float formulas, so a seed gives the reference's catalogue. The sampler
// (StarfieldGenerator::sampleStarfield, PHSF001A.h:153-201: a Gaussian point-spread sum over
// the catalogue) runs on the device: per star, the host precomputes what does not depend on
// the view direction (the direction as the reference normalises it, sigma^2 of the point-spread
// function, 10^(-0.4 m), the colour) and a cosine below which the spread is under its 1e-6 cut.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/sirius_hip.h"
#include "sirius_internal.hpp"

namespace sirius {

hipError_t launch_star_dirs(const float* rec, int n_star, float brightness, const float* dirs, int n, float* rgb,
                            hipStream_t st);

namespace {

// StarfieldConfig::validate (PHSF001A.h:112-121)
void validate(SiriusHipStarfield& c) {
    c.star_count = std::min(c.star_count, 10000000u);
    c.min_distance_pc = std::max(c.min_distance_pc, 0.1f);
    c.max_distance_pc = std::max(c.max_distance_pc, c.min_distance_pc + 1.0f);
    c.magnitude_limit = std::clamp(c.magnitude_limit, 0.0f, 20.0f);
    c.aperture_mm = std::clamp(c.aperture_mm, 0.0f, 1000.0f);
    c.focus_distance_pc = std::clamp(c.focus_distance_pc, c.min_distance_pc, c.max_distance_pc);
}

// One candidate star: five uniform draws (polar cosine, azimuth, volume-uniform distance,
// absolute magnitude on a Salpeter-like law, colour scatter), then the apparent magnitude,
// the B-V index and the temperature (PHSF001A.h:206-249).
SiriusHipStar draw_star(std::mt19937& rng, std::uniform_real_distribution<float>& U, const SiriusHipStarfield& c) {
    SiriusHipStar s;
    std::memset(&s, 0, sizeof(s));
    const float theta = std::acos(2.0f * U(rng) - 1.0f);
    const float phi = 2.0f * static_cast<float>(M_PI) * U(rng);
    s.dir[0] = std::sin(theta) * std::cos(phi);
    s.dir[1] = std::sin(theta) * std::sin(phi);
    s.dir[2] = std::cos(theta);
    const float u = U(rng);
    const float dmin = c.min_distance_pc, dmax = c.max_distance_pc;
    s.distance_pc = std::cbrt(u * (dmax * dmax * dmax) + (1.0f - u) * (dmin * dmin * dmin));
    const float abs_mag = -2.0f + 12.0f * std::pow(U(rng), 0.3f);
    s.magnitude = abs_mag + 5.0f * std::log10(s.distance_pc / 10.0f);
    const float t = (abs_mag + 2.0f) / 12.0f;
    s.color_bv = std::clamp(-0.3f + 2.0f * t + 0.3f * (U(rng) - 0.5f), -0.4f, 2.0f);
    const float temp = s.color_bv > -0.1f ? 4600.0f / (0.92f + s.color_bv) : 30000.0f + 10000.0f * (-s.color_bv - 0.1f);
    s.temperature_K = std::clamp(temp, 2500.0f, 50000.0f);
    return s;
}

// StarEntry::computeColor (PHSF001A.h:51-73): a Planckian-locus colour normalised to its largest
// channel.
void star_colour(float T, float& r, float& g, float& b) {
    if (T <= 0.0f) T = 5778.0f;
    const float tn = T / 6500.0f;
    if (T <= 6500.0f) {
        r = 1.0f;
        g = std::pow(tn, 0.4f);
        b = std::pow(tn, 0.8f);
    } else {
        r = std::pow(1.0f / tn, 0.3f);
        g = std::pow(1.0f / tn, 0.15f);
        b = 1.0f;
    }
    const float m = std::max({r, g, b});
    r /= m;
    g /= m;
    b /= m;
}

// Device record of a star (SIRIUS_STAR_REC floats): normalised direction, cosine cut, sigma^2,
// 10^(-0.4 m), colour, sine of the cut angle. The direction is normalised as sampleStarfield does for an observer at
// the origin (direction times distance, then divided by its length), in the same float steps.
void star_record(const SiriusHipStar& s, const SiriusHipStarfield& c, float* rec) {
    const float dx = s.dir[0] * s.distance_pc - 0.0f;
    const float dy = s.dir[1] * s.distance_pc - 0.0f;
    const float dz = s.dir[2] * s.distance_pc - 0.0f;
    const float d = std::sqrt(dx * dx + dy * dy + dz * dz);
    std::memset(rec, 0, SIRIUS_STAR_REC * sizeof(float));
    if (d < 1e-6f) {
        rec[3] = 2.0f;  // no direction: never within the cut
        return;
    }
    rec[0] = dx / d;
    rec[1] = dy / d;
    rec[2] = dz / d;
    // computePSFSigma (PHSF001A.h:252-269)
    const float base = 1e-6f;
    float dof = 0.0f;
    if (c.enable_dof && c.aperture_mm > 0.0f) {
        const float defocus = std::abs(1.0f / s.distance_pc - 1.0f / c.focus_distance_pc);
        dof = c.aperture_mm * defocus * 1e-3f;
    }
    const float sigma = std::sqrt(base * base + dof * dof);
    // exp(-a^2 / (2 sigma^2)) >= 1e-6 needs a <= sigma sqrt(2 ln 1e6); the cut is taken 1 % wider
    // (it only skips stars; the 1e-6 test itself runs on the device)
    const double amax = 1.01 * (double)sigma * std::sqrt(2.0 * std::log(1e6));
    rec[3] = amax >= M_PI ? -2.0f : (float)std::cos(amax);
    rec[9] = amax >= M_PI ? 0.0f : (float)std::sin(amax);  // with rec[3], widens the cut by a beam
    rec[4] = sigma * sigma;
    rec[5] = std::pow(10.0f, -0.4f * s.magnitude);  // StarEntry::intensity (PHSF001A.h:77-79)
    star_colour(s.temperature_K, rec[6], rec[7], rec[8]);
}

}  // namespace

}  // namespace sirius

using namespace sirius;

#define GUARD_BEGIN try {
#define GUARD_END(h)                                                                     \
    }                                                                                    \
    catch (const std::bad_alloc&) { return (h)->fail(SIRIUS_HIP_E_NOMEM, "host allocation"); } \
    catch (...) { return (h)->fail(SIRIUS_HIP_E_DEVICE, "unexpected exception"); }

extern "C" {

void sirius_hip_starfield_default(SiriusHipStarfield* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->star_count = 100000;
    c->seed = 42;
    c->min_distance_pc = 1.0f;
    c->max_distance_pc = 10000.0f;
    c->magnitude_limit = 12.0f;
    c->aperture_mm = 50.0f;
    c->focus_distance_pc = 100.0f;
    c->brightness_scale = 1.0f;
    c->enable_dof = 1;
}

int sirius_hip_starfield_generate(const SiriusHipStarfield* cfg, SiriusHipStar* stars, int capacity, int* count) {
    if (!cfg || !count || capacity < 0 || (capacity > 0 && !stars)) return SIRIUS_HIP_E_INVALID;
    SiriusHipStarfield c = *cfg;
    validate(c);
    std::mt19937 rng(c.seed);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    int n = 0;
    for (uint32_t i = 0; i < c.star_count; ++i) {
        SiriusHipStar s = draw_star(rng, U, c);
        if (s.magnitude <= c.magnitude_limit) {
            if (n < capacity) stars[n] = s;
            ++n;
        }
    }
    *count = n;
    return SIRIUS_HIP_OK;
}

int sirius_hip_set_starfield(SiriusHipHandle hh, int side, const SiriusHipStarfield* cfg, const SiriusHipStar* stars,
                             int n) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
        if (side < 0 || side > 1) return H->fail(SIRIUS_HIP_E_INVALID, "starfield side must be 0 or 1");
        if (n < 0 || (n > 0 && (!stars || !cfg))) return H->fail(SIRIUS_HIP_E_INVALID, "starfield arguments");
        int rc = H->ensure_device();
        if (rc) return rc;
        if (H->star_dev[side]) (void)hipFree(H->star_dev[side]);
        H->star_dev[side] = nullptr;
        H->star_n[side] = 0;
        if (n == 0) return SIRIUS_HIP_OK;
        SiriusHipStarfield c = *cfg;
        validate(c);
        std::vector<float> rec((size_t)n * SIRIUS_STAR_REC);
        for (int i = 0; i < n; ++i) star_record(stars[i], c, rec.data() + (size_t)i * SIRIUS_STAR_REC);
        hipError_t e = hipMalloc((void**)&H->star_dev[side], rec.size() * sizeof(float));
        if (e != hipSuccess) {
            H->star_dev[side] = nullptr;
            return H->hip_fail(e, "starfield allocation");
        }
        e = hipMemcpyAsync(H->star_dev[side], rec.data(), rec.size() * sizeof(float), hipMemcpyHostToDevice, H->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(H->stream);
        if (e != hipSuccess) return H->hip_fail(e, "starfield upload");
        H->star_n[side] = n;
        H->star_brightness[side] = c.brightness_scale;
        return SIRIUS_HIP_OK;
    GUARD_END(H)
}

int sirius_hip_sample_starfield(SiriusHipHandle hh, int side, const float* dirs, int n, float* rgb) {
    if (!hh) return SIRIUS_HIP_E_INVALID;
    Handle* H = static_cast<Handle*>(hh);
    GUARD_BEGIN
        if (side < 0 || side > 1 || n < 0 || (n > 0 && (!dirs || !rgb)))
            return H->fail(SIRIUS_HIP_E_INVALID, "sample_starfield arguments");
        if (n == 0) return SIRIUS_HIP_OK;
        int rc = H->ensure_device();
        if (rc) return rc;
        float* dd = (float*)H->scratch("star_dirs", sizeof(float) * 3 * (size_t)n, &rc);
        if (rc) return rc;
        float* dr = (float*)H->scratch("star_rgb", sizeof(float) * 3 * (size_t)n, &rc);
        if (rc) return rc;
        hipError_t e = hipMemcpyAsync(dd, dirs, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, H->stream);
        if (e == hipSuccess)
            e = launch_star_dirs(H->star_dev[side], H->star_n[side], H->star_brightness[side], dd, n, dr, H->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(rgb, dr, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost, H->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(H->stream);
        if (e != hipSuccess) return H->hip_fail(e, "sample_starfield");
        return SIRIUS_HIP_OK;
    GUARD_END(H)
}

}  // extern "C"
