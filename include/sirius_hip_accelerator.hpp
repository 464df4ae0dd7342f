// sirius_hip_accelerator.hpp — C++ IAccelerator backend over the sirius_hip C-ABI.
//
// Drop-in for the OptiX backend's wrapper (src/Sirius.Render/Acceleration/OptiX/RDOX001A.h:12-44):
// the same IAccelerator method set (src/Sirius.Render/Acceleration/Backend/ACIB001A.h:148-189),
// reporting BackendType::HIP (ACIB001A.h:28). Inside the Sirius tree define
// SIRIUS_HIP_IN_SIRIUS_TREE and the class derives from the tree's own IAccelerator; outside
// it (this repository's tests) a declaration-compatible interface is supplied below.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "sirius_hip.h"

#ifdef SIRIUS_HIP_IN_SIRIUS_TREE
#include "ACIB001A.h"
#else
namespace Sirius::Acceleration {

// Same enumerators and order as ACIB001A.h:24-32.
enum class BackendType : uint8_t { None, OptiX, CUDA, HIP, OneAPI, Metal, CPU };

// Same fields as ACIB001A.h:50-59.
struct DeviceCapabilities {
    std::string deviceName;
    size_t totalMemory = 0;
    size_t freeMemory = 0;
    int computeCapability = 0;
    bool supportsRTCores = false;
    bool supportsFP64 = false;
    int multiprocessorCount = 0;
    int maxThreadsPerBlock = 0;
};

// The LaunchConfig fields this backend reads (ACIB001A.h:64-96), same names and defaults.
struct LaunchConfig {
    int width = 1920;
    int height = 1080;
    int samplesPerPixel = 64;
    int tileSize = 64;
    float blackHoleMass = 1.0f;
    float blackHoleSpin = 0.0f;
    float observerDistance = 100.0f;
    float observerInclination = 1.5708f;
    float observerAzimuth = 0.0f;
    float cameraFOV = 45.0f;
    float diskInnerRadius = 6.0f;
    float diskOuterRadius = 30.0f;
    float diskTemperature = 50000.0f;
    float diskEmission = 2.5f;
    int temperatureModel = 1;
    float maxStepSize = 0.5f;
    int maxSteps = 10000;
    int metricType = 2;
    int metricFamily = 0;
    float throatRadius = 1.0f;
    float warpVelocity = 0.5f;
    float bubbleRadius = 1.0f;
    float bubbleSigma = 0.5f;
    bool enableStarfield = false;      // ACIB001A.h:110-112
    float starfieldBrightness = 1.0f;
};

// Same pure-virtual method set as ACIB001A.h:148-189.
class IAccelerator {
public:
    virtual ~IAccelerator() = default;
    virtual BackendType getType() const = 0;
    virtual DeviceCapabilities getCapabilities() const = 0;
    virtual bool initialise(int width, int height) = 0;
    virtual bool isInitialised() const = 0;
    virtual void launch(const LaunchConfig& config) = 0;
    virtual float* getFrameBuffer() = 0;
    virtual size_t getFrameBufferSize() const = 0;
    virtual bool uploadBackground(const uint8_t* data, int width, int height) = 0;
    virtual void synchronise() = 0;
    virtual void cleanup() = 0;
    virtual std::string getLastError() const = 0;
    virtual void resetAccumulation() {}
};

}  // namespace Sirius::Acceleration
#endif

namespace Sirius::Acceleration::HIP {

class HIPAccelerator final : public IAccelerator {
public:
    HIPAccelerator();
    ~HIPAccelerator() override;
    HIPAccelerator(const HIPAccelerator&) = delete;
    HIPAccelerator& operator=(const HIPAccelerator&) = delete;

    BackendType getType() const override { return BackendType::HIP; }
    DeviceCapabilities getCapabilities() const override;
    bool initialise(int width, int height) override;
    bool isInitialised() const override;
    void launch(const LaunchConfig& config) override;
    float* getFrameBuffer() override;
    size_t getFrameBufferSize() const override;
    bool uploadBackground(const uint8_t* data, int width, int height) override;
    void synchronise() override;
    void cleanup() override;
    std::string getLastError() const override;

    // Backend-specific: the batched physics entry point and the precision mode used by launch.
    int traceBatch(const SiriusHipCamera* cams, int n_cam, const SiriusHipMetric* metrics, int n_metric,
                   const SiriusHipTraceOpts& opts, SiriusHipRayOutput& out, SiriusHipStats* stats);
    void setPrecision(int precision) { m_precision = precision; }
    // Star light averaged over each pixel's ray bundle in later launches (point-sampled when off).
    bool setRayBundle(bool enabled);
    // Writes the framebuffer of the last launch (SIRIUS_IMAGE_*; exposure applies to the PPM),
    // converted on the device: what SREP001A.cpp:252-276 does with EXRWriter / BufferWriter.
    bool writeImage(const std::string& path, int format, float exposure = 1.0f);

    // LaunchConfig -> C-ABI descriptors (units of the hole's mass); the spacetime metricType names.
    // Returns SIRIUS_HIP_OK, or an error code with the reason in *why.
    static int convertConfig(const LaunchConfig& cfg, SiriusHipCamera& cam, SiriusHipMetric& metric,
                             SiriusHipTraceOpts& opts, int precision, std::string* why = nullptr);
    // LaunchConfig -> framebuffer shading (samplesPerPixel, temperatureModel, diskTemperature).
    static void convertShading(const LaunchConfig& cfg, SiriusHipShading& sh);

private:
    // LaunchConfig::enableStarfield -> the default catalogue on both celestial spheres
    void syncStarfield(const LaunchConfig& config);

    SiriusHipHandle m_handle = nullptr;
    int m_precision = SIRIUS_PRECISION_FP32_ESCALATE;
    std::string m_error;
    bool m_stars = false;
    float m_starBrightness = 0.0f;
    // launch shape from sirius_hip_autotune, kept per frame size and spacetime
    bool m_tuned = false;
    double m_tuneKey[7] = {};
    int m_tune[3] = {0, 0, 0};
};

// Factory hook for ACBF001A.cpp's createAccelerator(BackendType::HIP).
std::unique_ptr<IAccelerator> createHIPAccelerator();

}  // namespace Sirius::Acceleration::HIP
