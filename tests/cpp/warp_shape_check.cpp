// Host build of the Alcubierre shape function of sirius_amd/csrc/geodesic.hpp (warp_shape: one
// exponential of -2 sigma (r - R), by tanh A - tanh B = sinh(A - B) / (cosh A cosh B)) against its
// definition
// f = (tanh(sigma (r + R)) - tanh(sigma (r - R))) / (2 tanh(sigma R)) and the derivative of that
// definition, both in long double (the difference of tanh and 1 - tanh^2 written without
// cancellation, and the plain difference checked where it keeps its digits), on seeded random
// (sigma, R, r). Bars: FP64 within 2e-13 relative (the exponential's argument, up to 400, carries
// its rounding into v at about 4e-14; df/dr relative to 1e-3 sigma where it vanishes near r = 0);
// FP32 within 1.5 times the error of the two-tanh definition evaluated in FP32. Prints "OK" when
// every case meets them.
#include "../../sirius_amd/csrc/geodesic.hpp"

#include <cmath>
#include <cstdio>
#include <random>

using namespace sirius;

template <class T>
static Spacetime<T> warp(double sigma, double R) {
    Spacetime<T> s{};
    s.family = 2;
    s.wv = T(0.5);
    s.wR = T(R);
    s.wsig = T(sigma);
    s.winv2tanh = T(1.0 / (2.0 * std::tanh(sigma * R)));
    s.wE = T(std::exp(-4.0 * sigma * R));
    s.wN = T((1.0 - std::exp(-4.0 * sigma * R)) / std::tanh(sigma * R));
    return s;
}

int main() {
    std::mt19937_64 g(0x3a11);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    double worst64 = 0.0, worst32 = 0.0, worst32_tanh = 0.0;
    for (int t = 0; t < 200000; ++t) {
        const double sigma = 0.5 + 9.5 * U(g), R = 0.5 + 2.5 * U(g), r = 20.0 * U(g) * U(g);
        const long double A = (long double)sigma * (r + R), B = (long double)sigma * (r - R);
        const long double n = 1.0L / (2.0L * std::tanh((long double)sigma * R));
        // the definition, with tanh A - tanh B and 1 - tanh^2 written without cancellation:
        // sinh(A - B) / (cosh A cosh B) and 1 / cosh^2
        const long double cA = std::cosh(A), cB = std::cosh(B);
        const long double f = std::sinh(A - B) / (cA * cB) * n;
        const long double df = (long double)sigma * (1.0L / (cA * cA) - 1.0L / (cB * cB)) * n;
        // and the plain difference of tanh where it keeps its digits
        if (std::fabs(f) > 1e-3L) {
            const long double fd = (std::tanh(A) - std::tanh(B)) * n;
            worst64 = std::fmax(worst64, (double)(std::fabs(fd - f) / std::fabs(f)) * 1e-13 / 1e-15);
        }
        double d64;
        const double f64 = warp_shape(warp<double>(sigma, R), r, d64);
        float d32;
        const float f32 = warp_shape(warp<float>(sigma, R), (float)r, d32);
        const long double ef = std::fabs(f64 - f) / (std::fabs(f) + 1e-300L);
        // df/dr vanishes at r = 0, where 1 - u^2 cancels: relative to its scale sigma there
        const long double ed = std::fabs(d64 - df) / std::fmax(std::fabs(df), 1e-3L * sigma);
        worst64 = std::fmax(worst64, (double)std::fmax(ef, ed));
        {
            // the definition itself in FP32 (two tanh), for the FP32 bar
            const float tp = std::tanh((float)sigma * ((float)r + (float)R)), tm = std::tanh((float)sigma * ((float)r - (float)R));
            const float nn = (float)(1.0 / (2.0 * std::tanh(sigma * R)));
            const float ft = (tp - tm) * nn, dt = (float)sigma * ((1.f - tp * tp) - (1.f - tm * tm)) * nn;
            worst32_tanh = std::fmax(worst32_tanh, (double)std::fmax(std::fabs(ft - f), std::fabs(dt - df) / (std::fabs(df) + 1.0L)));
        }
        worst32 = std::fmax(worst32, (double)std::fmax(std::fabs(f32 - f), std::fabs(d32 - df) / (std::fabs(df) + 1.0L)));
    }
    std::printf("FP64 worst %.3e, FP32 worst %.3e (the two-tanh form in FP32: %.3e)\n", worst64, worst32, worst32_tanh);
    const bool ok = worst64 < 2e-13 && worst32 < 1.5 * worst32_tanh;
    std::printf("%s\n", ok ? "OK" : "FAIL");
    return ok ? 0 : 1;
}
