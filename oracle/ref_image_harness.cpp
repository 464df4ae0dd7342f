// ref_image_harness.cpp — drives the reference's own image writers (compiled in place from the
// reference tree by oracle/Makefile, target `ref`) to write golden image files, and the
// decoders vendored with the reference (stb_image, tinyexr) to read files back.
// TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py and bench.py's CPU-baseline
// leg; never linked into the product.
//
// What is the reference's own code here (called, not restated):
//   PNGWriter::writeRGBA             Sirius.Render/Output/OUPN001A.cpp:51-72 (stb_image_write)
//   EXRWriter::writeEXR (RGBA)       Sirius.Render/Output/OUEW001A.cpp:109-178,181-190 (tinyexr)
//   EXRWriter::writePPM / writePFM   Sirius.Render/Output/OUEW001A.h:97-126
//   BufferWriter::writePPM / writeRaw  Sirius.Render/Buffer/BFIO001A.h:44-86
//   stbi_load_from_memory            lib/stb/stb_image.h
//   LoadEXRFromMemory                lib/tinyexr/tinyexr.h
#include "Output/OUEW001A.h"
#include "Output/OUPN001A.h"
#include "Buffer/BFIO001A.h"

#define STB_IMAGE_IMPLEMENTATION
#include "stb/stb_image.h"
#include "tinyexr.h"

#include <cstdint>
#include <cstdlib>
#include <cstring>

using sirius::render::EXRWriter;
using sirius::render::ImageBuffer;
using sirius::render::ImageBufferRGBA;
using sirius::render::PNGWriter;

namespace {

ImageBuffer rgb_buffer(int w, int h, const float* rgb) {
    ImageBuffer b;
    b.allocate(w, h);
    std::memcpy(b.pixels.data(), rgb, sizeof(float) * 3 * (size_t)w * h);
    return b;
}

}  // namespace

extern "C" {

int ref_write_png_rgba(const char* path, int w, int h, const float* rgba) {
    return PNGWriter::writeRGBA(path, w, h, rgba) ? 1 : 0;
}

int ref_write_exr_rgba(const char* path, int w, int h, const float* rgba) {
    ImageBufferRGBA b;
    b.allocate(w, h);
    std::memcpy(b.pixels.data(), rgba, sizeof(float) * 4 * (size_t)w * h);
    return EXRWriter::writeEXR(path, b) ? 1 : 0;
}

int ref_write_ppm_aces(const char* path, const float* rgba, int w, int h, float exposure) {
    return Sirius::Buffer::BufferWriter::writePPM(path, rgba, w, h, exposure) ? 1 : 0;
}

int ref_write_ppm_srgb(const char* path, int w, int h, const float* rgb) {
    return EXRWriter::writePPM(path, rgb_buffer(w, h, rgb)) ? 1 : 0;
}

int ref_write_pfm(const char* path, int w, int h, const float* rgb) {
    return EXRWriter::writePFM(path, rgb_buffer(w, h, rgb)) ? 1 : 0;
}

int ref_write_raw(const char* path, const float* data, int w, int h, int channels) {
    return Sirius::Buffer::BufferWriter::writeRaw(path, data, w, h, channels) ? 1 : 0;
}

// Decodes a PNG held in memory to 8-bit RGBA; returns 1 and the size, or 0.
int ref_load_png(const unsigned char* bytes, int n, int* w, int* h, unsigned char* out, int capacity) {
    int c = 0;
    unsigned char* px = stbi_load_from_memory(bytes, n, w, h, &c, 4);
    if (!px) return 0;
    const size_t sz = (size_t)(*w) * (*h) * 4;
    if (sz <= (size_t)capacity) std::memcpy(out, px, sz);
    stbi_image_free(px);
    return sz <= (size_t)capacity ? 1 : 0;
}

// Decodes an EXR held in memory to float RGBA (tinyexr's own half -> float); returns 1 or 0.
int ref_load_exr(const unsigned char* bytes, int n, int* w, int* h, float* out, int capacity) {
    float* px = nullptr;
    const char* err = nullptr;
    if (LoadEXRFromMemory(&px, w, h, bytes, (size_t)n, &err) != TINYEXR_SUCCESS) {
        if (err) FreeEXRErrorMessage(err);
        return 0;
    }
    const size_t sz = (size_t)(*w) * (*h) * 4;
    if (sz <= (size_t)capacity) std::memcpy(out, px, sz * sizeof(float));
    std::free(px);
    return sz <= (size_t)capacity ? 1 : 0;
}

}  // extern "C"
