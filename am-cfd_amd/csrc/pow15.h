// v^1.5 for the free-surface energy balance (FreeSurfaceManager.cpp:301,
// where the reference calls pow(v, 1.5) with v = 1 + |grad z|^2 >= 1).
//
// glibc's pow is not reproducible from first principles (its logarithm
// table is a rounding choice of the library), but it is correctly rounded
// except when the exact result lies extremely close to a rounding boundary.
// This evaluates v * sqrt(v) with the rounding errors of the square root and
// of the product carried in a second double (exact FMA residuals), so the
// result is the correctly rounded v^1.5 except in the same kind of
// near-boundary cases; tests/native/pow15_check.cpp measures the agreement
// with the C library.
#pragma once

#include <hip/hip_runtime.h>

namespace amcfd {

__host__ __device__ inline double pow15(double v) {
    const double s = __builtin_sqrt(v);          // correctly rounded sqrt(v)
    const double rem = __builtin_fma(-s, s, v);  // v - s^2, exact
    const double ds = rem / (2.0 * s);           // sqrt(v) - s to first order
    const double p = v * s;
    const double perr = __builtin_fma(v, s, -p); // v*s - p, exact
    return p + (perr + v * ds);
}

}  // namespace amcfd
