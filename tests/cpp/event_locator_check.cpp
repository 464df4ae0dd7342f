// Host build of the tracer's event locators (sirius_amd/csrc/geodesic.hpp): the safeguarded
// Newton iteration the FP32 path uses against a 52-halving bisection of the same dense-output
// polynomial, on seeded random steps that cross the plane z = 0 or a sphere of the Kerr radius
// (a = 0.9) once. A step is a quartic curve x(t) = c0 + c1 t + c2 t^2 + c3 t^3 + c4 t^4 over
// [0, h], which the dense output represents exactly (ends, end slopes and the t^4 coefficient
// c4 h^4 as its fifth coefficient). Bars: the event function (in FP64, on the FP32 coefficients)
// at Newton's th is within twice its value at an FP32 bisection's th plus 1e-6 (1 + target), a
// few FP32 rounding units of the position; where the crossing is well conditioned (slope of the
// event function above 0.5 per unit th) th is within 1e-5 of the FP64 bisection's. Prints the
// worst cases and "OK" when every case meets both.
#include "../../sirius_amd/csrc/geodesic.hpp"

#include <cmath>
#include <cstdio>
#include <random>

using namespace sirius;

// bisection of the event function on the dense output, in FP64 from the FP32 coefficients
static void widen(const float x0[4], const float k0[4], const float x1[4], const float k1[4], const float dn[6],
                  double X0[4], double K0[4], double X1[4], double K1[4], double DN[6]) {
    for (int i = 0; i < 4; ++i) { X0[i] = x0[i]; K0[i] = k0[i]; X1[i] = x1[i]; K1[i] = k1[i]; }
    for (int i = 0; i < 6; ++i) DN[i] = dn[i];
}

// the event function and its slope dth on the dense output at th, in FP64
static double event_at(const Spacetime<double>& s, int kind, double target, const float x0[4], const float k0[4],
                       const float x1[4], const float k1[4], const float dn[6], float h, double th, double* slope) {
    double X0[4], K0[4], X1[4], K1[4], DN[6], xm[4], v[4], rate;
    widen(x0, k0, x1, k1, dn, X0, K0, X1, K1, DN);
    dense_pos(X0, K0, X1, K1, DN, (double)h, th, xm);
    for (int i = 1; i < 4; ++i) v[i] = dense_slope(X0[i], X1[i], (double)h, K0[i], K1[i], DN[i - 1], th);
    if (kind == EV_Z) {
        *slope = v[3];
        return xm[3];
    }
    const double r = event_radius_rate(s, xm, v, rate);
    *slope = rate;
    return r - target;
}

static double bisect(const Spacetime<double>& s, int kind, double target, const float x0[4], const float k0[4],
                     const float x1[4], const float k1[4], const float dn[6], float h) {
    double X0[4], K0[4], X1[4], K1[4], DN[6];
    widen(x0, k0, x1, k1, dn, X0, K0, X1, K1, DN);
    return locate_event_dense(s, kind, target, X0, K0, X1, K1, DN, (double)h);
}

int main() {
    std::mt19937_64 g(0x5eed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    Spacetime<float> sf{};
    Spacetime<double> sd{};
    sf.family = sd.family = 0;
    sf.M = 1.f; sd.M = 1.0;
    sf.a = 0.9f; sd.a = (double)0.9f;
    sf.a2 = sf.a * sf.a; sd.a2 = sd.a * sd.a;
    double worst[2] = {0.0, 0.0};  // worst excess over the bars, plane and sphere
    int n_z = 0, n_r = 0, fails = 0;
    // Newton's root against the bisections' on one crossing; returns the excess over the bars
    auto check = [&](int kind, float target, const float x0[4], const float k0[4], const float x1[4],
                     const float k1[4], const float dn[6], float h, float tn) {
        const double tb = bisect(sd, kind, (double)target, x0, k0, x1, k1, dn, h);
        const float t32 = locate_event_dense(sf, kind, target, x0, k0, x1, k1, dn, h);
        double sl, sl32, slb;
        const double fn = std::fabs(event_at(sd, kind, target, x0, k0, x1, k1, dn, h, tn, &sl));
        const double f32 = std::fabs(event_at(sd, kind, target, x0, k0, x1, k1, dn, h, t32, &sl32));
        event_at(sd, kind, target, x0, k0, x1, k1, dn, h, tb, &slb);
        const double bar_f = 2.0 * f32 + 1e-6 * (1.0 + std::fabs(target));
        double excess = fn / bar_f;
        if (std::fabs(slb) > 0.5) excess = std::fmax(excess, std::fabs(tn - tb) / 1e-5);
        if (excess > 1.0 && fails++ < 5)
            std::printf("kind %d th %.7f bisection %.7f (FP32 %.7f) residual %.3e (FP32 bisection %.3e) slope %.3e\n",
                        kind, tn, tb, t32, fn, f32, slb);
        return excess;
    };
    for (int t = 0; t < 4000000 && (n_z < 50000 || n_r < 50000); ++t) {
        const bool plane = n_z < 50000 && (n_r >= 50000 || (t & 1));
        float x0[4], k0[4], x1[4], k1[4], dn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const double h = 0.2 + 3.0 * std::abs(U(g));
        x0[0] = x1[0] = 0.f;
        k0[0] = k1[0] = 1.f;
        for (int i = 1; i < 4; ++i) {
            const double c0 = 8.0 * U(g), c1 = U(g), c2 = 0.05 * U(g), c3 = 0.005 * U(g), c4 = 0.0005 * U(g);
            x0[i] = (float)c0;
            k0[i] = (float)c1;
            x1[i] = (float)(c0 + h * (c1 + h * (c2 + h * (c3 + h * c4))));
            k1[i] = (float)(c1 + h * (2.0 * c2 + h * (3.0 * c3 + h * 4.0 * c4)));
            dn[i - 1] = (float)(c4 * h * h * h * h);
        }
        // one crossing only: with several, both locators return a root, not necessarily the same
        auto crossings = [&](int kind, float target) {
            int n = 0;
            float xm[4];
            double prev = 0.0;
            for (int j = 0; j <= 256; ++j) {
                dense_pos(x0, k0, x1, k1, dn, (float)h, j / 256.f, xm);
                const double v = event_value(sf, kind, xm, target);
                if (j > 0 && (v < 0.0) != (prev < 0.0)) ++n;
                prev = v;
            }
            return n;
        };
        if (plane) {
            if (x0[3] * x1[3] >= 0.f || crossings(EV_Z, 0.f) != 1) continue;
            const float tn = locate_event_newton(sf, EV_Z, 0.f, x0, k0, x1, k1, dn, (float)h, x0[3], x1[3]);
            worst[0] = std::fmax(worst[0], check(EV_Z, 0.f, x0, k0, x1, k1, dn, (float)h, tn));
            ++n_z;
        } else {
            const float r0 = kerr_r(sf, x0[1], x0[2], x0[3]), r1 = kerr_r(sf, x1[1], x1[2], x1[3]);
            const float target = 0.5f * (r0 + r1) + (float)(0.3 * U(g)) * std::fabs(r1 - r0);
            if ((r0 - target) * (r1 - target) >= 0.f || std::fmin(r0, r1) < 2.f || crossings(EV_R, target) != 1)
                continue;
            const float tn =
                locate_event_newton(sf, EV_R, target, x0, k0, x1, k1, dn, (float)h, r0 - target, r1 - target);
            worst[1] = std::fmax(worst[1], check(EV_R, target, x0, k0, x1, k1, dn, (float)h, tn));
            ++n_r;
        }
    }
    std::printf("plane crossings %d, worst fraction of the bars %.3f\n", n_z, worst[0]);
    std::printf("sphere crossings %d, worst fraction of the bars %.3f\n", n_r, worst[1]);
    const bool ok = n_z >= 50000 && n_r >= 50000 && worst[0] <= 1.0 && worst[1] <= 1.0;
    std::printf("%s\n", ok ? "OK" : "FAIL");
    return ok ? 0 : 1;
}
