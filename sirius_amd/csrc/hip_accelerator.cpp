// hip_accelerator.cpp — HIPAccelerator: the IAccelerator implementation over the C-ABI.
// Mirrors the OptiX wrapper's shape (src/Sirius.Render/Acceleration/OptiX/RDOX001A.cpp:91-341:
// launch -> convertConfig -> C API launch; getFrameBuffer returns the C layer's host copy).
#include "../../include/sirius_hip_accelerator.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace Sirius::Acceleration::HIP {

HIPAccelerator::HIPAccelerator() : m_handle(sirius_hip_create()) {
    if (!m_handle) m_error = "sirius_hip_create failed";
}

HIPAccelerator::~HIPAccelerator() {
    if (m_handle) sirius_hip_destroy(m_handle);
}

DeviceCapabilities HIPAccelerator::getCapabilities() const {
    DeviceCapabilities caps;
    if (!m_handle) return caps;
    char name[256] = {0};
    size_t total = 0, free_b = 0;
    int cus = 0, mtpb = 0;
    if (sirius_hip_device_info(m_handle, name, sizeof(name), &total, &free_b, &cus, &mtpb) == SIRIUS_HIP_OK) {
        caps.deviceName = name;
        caps.totalMemory = total;
        caps.freeMemory = free_b;
        caps.computeCapability = 0;  // CUDA-only notion
        caps.supportsRTCores = false;
        caps.supportsFP64 = true;
        caps.multiprocessorCount = cus;
        caps.maxThreadsPerBlock = mtpb;
    }
    return caps;
}

bool HIPAccelerator::initialise(int width, int height) {
    if (!m_handle) return false;
    int rc = sirius_hip_initialize(m_handle, width, height);
    if (rc != SIRIUS_HIP_OK) {
        m_error = sirius_hip_last_error(m_handle);
        return false;
    }
    return true;
}

bool HIPAccelerator::isInitialised() const { return m_handle && sirius_hip_is_initialized(m_handle); }

int HIPAccelerator::convertConfig(const LaunchConfig& cfg, SiriusHipCamera& cam, SiriusHipMetric& metric,
                                  SiriusHipTraceOpts& opts, int precision, std::string* why) {
    // Lengths in units of M (G = c = M = 1); blackHoleSpin is a/M. The spacetime is the one
    // metricType names (MetricType, RDOP003A.h:63-75: the OptiX backend selects its ray-generation
    // program by it, RDOX001A.cpp:102); metricFamily must agree with it.
    double M = cfg.blackHoleMass > 0 ? cfg.blackHoleMass : 1.0;
    int family;
    double a = std::clamp((double)cfg.blackHoleSpin, -0.9999, 0.9999);
    switch (cfg.metricType) {
        case 1:  // Schwarzschild: Kerr-Schild with a = Q = 0
        case 3:  // Reissner-Nordstrom: a = 0 (LaunchConfig carries no charge: Q = 0)
            family = SIRIUS_METRIC_KERR_SCHILD;
            a = 0.0;
            break;
        case 2:  // Kerr
        case 6:  // Kerr-Schild, native
            family = SIRIUS_METRIC_KERR_SCHILD;
            break;
        case 7:  // Ellis drainhole: the Morris-Thorne family
            family = SIRIUS_METRIC_MORRIS_THORNE;
            break;
        case 8:  // Alcubierre: the warp-drive family
            family = SIRIUS_METRIC_ALCUBIERRE;
            break;
        default:
            if (why)
                *why = "metricType " + std::to_string(cfg.metricType) +
                       " (Minkowski 0, Godel 4, Taub-NUT 5, de Sitter 9, custom 99) is not supported by the HIP backend";
            return SIRIUS_HIP_E_UNSUPPORTED;
    }
    if (cfg.metricFamily != family) {
        if (why)
            *why = "metricType " + std::to_string(cfg.metricType) + " needs metricFamily " + std::to_string(family) +
                   ", not " + std::to_string(cfg.metricFamily);
        return SIRIUS_HIP_E_INVALID;
    }
    std::memset(&metric, 0, sizeof(metric));
    metric.family = family;
    if (family == SIRIUS_METRIC_MORRIS_THORNE) {
        metric.throat_b0 = cfg.throatRadius / M;
    } else if (family == SIRIUS_METRIC_ALCUBIERRE) {
        metric.warp_v = cfg.warpVelocity;
        metric.warp_R = cfg.bubbleRadius / M;
        metric.warp_sigma = cfg.bubbleSigma * M;
    } else {
        metric.M = 1.0;
        metric.a = a;
    }
    std::memset(&cam, 0, sizeof(cam));
    cam.r_obs = cfg.observerDistance / M;
    cam.theta_obs = cfg.observerInclination;
    cam.phi_obs = cfg.observerAzimuth;
    cam.fov_deg = cfg.cameraFOV;
    cam.disk_r_in = cfg.diskInnerRadius / M;
    cam.disk_r_out = cfg.diskOuterRadius / M;
    cam.width = cfg.width;
    cam.height = cfg.height;
    cam.metric_index = 0;
    sirius_hip_default_opts(&opts, precision);
    opts.n_bins = 16;
    if (cfg.maxSteps > 0) opts.max_steps = cfg.maxSteps;
    return SIRIUS_HIP_OK;
}

void HIPAccelerator::convertShading(const LaunchConfig& cfg, SiriusHipShading& sh) {
    sirius_hip_default_shading(&sh);
    sh.samples_per_pixel = std::clamp(cfg.samplesPerPixel, 1, 4096);
    sh.temperature_model = cfg.temperatureModel == 0 ? 0 : 1;
    sh.disk_temperature = std::max(0.0, (double)cfg.diskTemperature);
}

// The OptiX backend turns LaunchConfig::enableStarfield into its kernel's procedural star field
// with starfieldBrightness (RDOX001A.cpp:269-274). Here it installs the reference's default
// catalogue (StarfieldConfig defaults, seed 42) as the sky, and the next seed's as the far side
// of a wormhole throat; both are regenerated only when the settings change.
void HIPAccelerator::syncStarfield(const LaunchConfig& config) {
    if (!config.enableStarfield) {
        if (m_stars) {
            sirius_hip_set_starfield(m_handle, 0, nullptr, nullptr, 0);
            sirius_hip_set_starfield(m_handle, 1, nullptr, nullptr, 0);
            m_stars = false;
        }
        return;
    }
    if (m_stars && m_starBrightness == config.starfieldBrightness) return;
    SiriusHipStarfield sf;
    sirius_hip_starfield_default(&sf);
    sf.brightness_scale = config.starfieldBrightness;
    for (int side = 0; side < 2; ++side) {
        sf.seed = 42u + (uint32_t)side;
        int n = 0;
        if (sirius_hip_starfield_generate(&sf, nullptr, 0, &n) != SIRIUS_HIP_OK) return;
        std::vector<SiriusHipStar> stars((size_t)std::max(n, 1));
        if (sirius_hip_starfield_generate(&sf, stars.data(), n, &n) != SIRIUS_HIP_OK ||
            sirius_hip_set_starfield(m_handle, side, &sf, stars.data(), n) != SIRIUS_HIP_OK) {
            m_error = sirius_hip_last_error(m_handle);
            return;
        }
    }
    m_stars = true;
    m_starBrightness = config.starfieldBrightness;
}

bool HIPAccelerator::setRayBundle(bool enabled) {
    if (!m_handle) return false;
    SiriusHipRayBundle b;
    sirius_hip_default_ray_bundle(&b);
    b.enabled = enabled ? 1 : 0;
    if (sirius_hip_set_ray_bundle(m_handle, &b) != SIRIUS_HIP_OK) {
        m_error = sirius_hip_last_error(m_handle);
        return false;
    }
    return true;
}

bool HIPAccelerator::writeImage(const std::string& path, int format, float exposure) {
    if (!m_handle) return false;
    SiriusHipImageOpts io;
    sirius_hip_default_image_opts(&io, format);
    io.exposure = exposure;
    if (sirius_hip_write_image(m_handle, path.c_str(), nullptr, 0, 0, 0, &io, nullptr) != SIRIUS_HIP_OK) {
        m_error = sirius_hip_last_error(m_handle);
        return false;
    }
    return true;
}

void HIPAccelerator::launch(const LaunchConfig& config) {
    if (!m_handle) return;
    syncStarfield(config);
    SiriusHipCamera cam;
    SiriusHipMetric metric;
    SiriusHipTraceOpts opts;
    if (convertConfig(config, cam, metric, opts, m_precision, &m_error) != SIRIUS_HIP_OK) return;
    SiriusHipShading shading;
    convertShading(config, shading);
    if (sirius_hip_set_shading(m_handle, &shading) != SIRIUS_HIP_OK) {
        m_error = sirius_hip_last_error(m_handle);
        return;
    }
    // The FP32 path's launch shape is tuned on the first frame of each size and spacetime (shapes
    // never change results); if tuning fails, the default shape is kept.
    if (opts.precision == SIRIUS_PRECISION_FP32_ESCALATE && opts.integrator == SIRIUS_INTEGRATOR_DP45) {
        const double key[7] = {double(cam.width), double(cam.height), double(metric.family), metric.M, metric.a,
                               metric.Q, metric.Lambda};
        bool same = m_tuned;
        for (int i = 0; i < 7; ++i) same = same && key[i] == m_tuneKey[i];
        if (!same) {
            SiriusHipTraceOpts t = opts;
            if (sirius_hip_autotune(m_handle, &cam, 1, &metric, 1, &t, nullptr) == SIRIUS_HIP_OK) {
                m_tune[0] = t.block_size;
                m_tune[1] = t.blocks_per_cu;
                m_tune[2] = t.refill_min;
            } else {
                m_tune[0] = m_tune[1] = m_tune[2] = 0;
            }
            m_tuned = true;
            for (int i = 0; i < 7; ++i) m_tuneKey[i] = key[i];
        }
        opts.block_size = m_tune[0];
        opts.blocks_per_cu = m_tune[1];
        opts.refill_min = m_tune[2];
    }
    int rc = sirius_hip_launch(m_handle, &cam, &metric, &opts);
    if (rc != SIRIUS_HIP_OK) m_error = sirius_hip_last_error(m_handle);
}

float* HIPAccelerator::getFrameBuffer() { return m_handle ? sirius_hip_get_frame_buffer(m_handle) : nullptr; }

size_t HIPAccelerator::getFrameBufferSize() const {
    return m_handle ? sirius_hip_get_frame_buffer_size(m_handle) : 0;
}

bool HIPAccelerator::uploadBackground(const uint8_t* data, int width, int height) {
    if (!m_handle) return false;
    int rc = sirius_hip_upload_background(m_handle, data, width, height);
    if (rc != SIRIUS_HIP_OK) {
        m_error = sirius_hip_last_error(m_handle);
        return false;
    }
    return true;
}

void HIPAccelerator::synchronise() {
    if (m_handle && sirius_hip_synchronize(m_handle) != SIRIUS_HIP_OK) m_error = sirius_hip_last_error(m_handle);
}

void HIPAccelerator::cleanup() {
    if (m_handle) sirius_hip_cleanup(m_handle);
}

std::string HIPAccelerator::getLastError() const { return m_error; }

int HIPAccelerator::traceBatch(const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics, int n_metric,
                               const SiriusHipTraceOpts& opts, SiriusHipRayOutput& out, SiriusHipStats* stats) {
    if (!m_handle) return SIRIUS_HIP_E_STATE;
    int rc = sirius_hip_trace_batch(m_handle, cams, n_cam, metrics, n_metric, &opts, &out, stats);
    if (rc != SIRIUS_HIP_OK) m_error = sirius_hip_last_error(m_handle);
    return rc;
}

std::unique_ptr<IAccelerator> createHIPAccelerator() { return std::make_unique<HIPAccelerator>(); }

}  // namespace Sirius::Acceleration::HIP
