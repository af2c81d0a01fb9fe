// The reference's own float64 arithmetic primitives (stereonet.py:29-205):
// square root by Newton iteration from a power-of-4 seed, e**x by range
// reduction plus a 15-term Taylor series, ln x by the atanh series, cos by
// range reduction plus an 11-term series.  The mesh and splatting code of
// the reference calls these instead of a math library, so results that must
// agree with it bit for bit evaluate the same steps in the same order
// (host and device; files including this are built without floating-point
// contraction).
#pragma once

#include <hip/hip_runtime.h>

namespace refm {

constexpr double kLn2 = 0.6931471805599453;

__host__ __device__ inline double fsqrt(double x) {
    if (x <= 0.0) return 0.0;
    double y = x;
    if (y > 4.0) {
        double s = 1.0;
        while (y > 4.0) {
            y *= 0.25;
            s *= 2.0;
        }
        y = s;
    } else if (y < 0.25) {
        double s = 1.0;
        while (y < 0.25) {
            y *= 4.0;
            s *= 0.5;
        }
        y = s;
    } else {
        y = 1.0;
    }
    for (int i = 0; i < 24; ++i) y = 0.5 * (y + x / y);
    return y;
}

__host__ __device__ inline double fexp(double x) {
    if (x > 700.0) return 1.0e308;
    if (x < -700.0) return 0.0;
    // k = int(x / LN2 +- 0.5), truncating toward zero
    const double kq = x / kLn2 + (x >= 0.0 ? 0.5 : -0.5);
    const long long k = (long long)kq;
    const double r = x - (double)k * kLn2;
    double term = 1.0, total = 1.0;
    for (int n = 1; n < 16; ++n) {
        term = term * r / (double)n;
        total += term;
    }
    if (k >= 0)
        for (long long i = 0; i < k; ++i) total *= 2.0;
    else
        for (long long i = 0; i < -k; ++i) total *= 0.5;
    return total;
}

__host__ __device__ inline double flog(double x) {
    if (x <= 0.0) return -1.0e308;
    int k = 0;
    while (x > 1.5) {
        x *= 0.5;
        ++k;
    }
    while (x < 0.5) {
        x *= 2.0;
        --k;
    }
    const double z = (x - 1.0) / (x + 1.0);
    const double z2 = z * z;
    double total = 0.0, term = z;
    for (int n = 1; n < 40; n += 2) {
        total += term / (double)n;
        term *= z2;
    }
    return 2.0 * total + (double)k * kLn2;
}

__host__ __device__ inline double fcos(double x) {
    const double two_pi = 6.283185307179586;
    while (x > 3.141592653589793) x -= two_pi;
    while (x < -3.141592653589793) x += two_pi;
    const double x2 = x * x;
    double term = 1.0, total = 1.0;
    for (int n = 1; n < 12; ++n) {
        term = term * (-x2) / (double)((2 * n - 1) * (2 * n));
        total += term;
    }
    return total;
}

__host__ __device__ inline double fabs_(double x) { return x < 0.0 ? -x : x; }
__host__ __device__ inline double fmax_(double a, double b) { return a > b ? a : b; }
__host__ __device__ inline double fmin_(double a, double b) { return a < b ? a : b; }

// sigmoid, stereonet.py:2539-2543
__host__ __device__ inline double sigmoid(double x) {
    if (x >= 0.0) return 1.0 / (1.0 + fexp(-x));
    const double e = fexp(x);
    return e / (1.0 + e);
}

}  // namespace refm
