// log() reproducing the C library's result bit for bit, for host and device.
//
// The reference's Fe-O surfactant model calls std::log (BoundCondManager.cpp:730),
// which on its x86-64 Linux hosts is glibc 2.35's log in the FMA ifunc
// variant.  That routine is the published table-driven algorithm of Szabolcs
// Nagy (ARM optimized-routines, 2018):
//   x = 2^k z, z in [0x1.6p-1, 0x1.6p0) split into 128 subintervals,
//   log(x) = k ln2 + log(c) + log1p(r),  r = z/c - 1 = fma(z, 1/c, -1),
//   log1p(r) ~= r + r^2 A(r) with a degree-6 polynomial,
// with the sum carried as hi + lo so that the result is accurate to about
// 0.52 ulp; inputs within [1 - 2^-4, 1 + 0x1.09p-4) take a degree-12
// polynomial of r = x - 1 with the r^2/2 term split exactly.  It is not
// correctly rounded, so a device log of any other construction disagrees with
// the host in the last bit often enough to perturb a SIMPLE run.  The table
// of {1/c, log c} is produced by gen_log_table.c from the published selection
// procedure; the polynomial coefficients and the multiply-adds the library's
// build fuses are kept.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "log_table.h"

namespace amcfd {

namespace logdetail {
__host__ __device__ inline uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
__host__ __device__ inline double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }

constexpr double kLn2hi = 0x1.62e42fefa3800p-1;   // ln 2, high part: k * kLn2hi is exact for |k| < 2^11
constexpr double kLn2lo = 0x1.ef35793c76730p-45;  // ln 2 - kLn2hi
// log1p(r) - r ~= r^2 (A0 + r A1 + r^2 A2 + r^3 A3 + r^4 A4) on |r| <= 0x1.fp-9
constexpr double kA0 = -0x1.0000000000001p-1, kA1 = 0x1.555555551305bp-2, kA2 = -0x1.fffffffeb459p-3;
constexpr double kA3 = 0x1.999b324f10111p-3, kA4 = -0x1.55575e506c89fp-3;
// log1p(r) ~= r + B0 r^2 + r^3 (B1 + r B2 + ...) on [-2^-4, 0x1.09p-4]
constexpr double kB0 = -0x1p-1, kB1 = 0x1.5555555555577p-2, kB2 = -0x1.ffffffffffdcbp-3;
constexpr double kB3 = 0x1.999999995dd0cp-3, kB4 = -0x1.55555556745a7p-3, kB5 = 0x1.24924a344de3p-3;
constexpr double kB6 = -0x1.fffffa4423d65p-4, kB7 = 0x1.c7184282ad6cap-4, kB8 = -0x1.999eb43b068ffp-4;
constexpr double kB9 = 0x1.78182f7afd085p-4, kB10 = -0x1.5521375d145cdp-4;
constexpr uint64_t kOff = 0x3fe6000000000000ull;  // bits of 0x1.6p-1, the start of the reduced range

__host__ __device__ inline double near_one(double x) {
    if (bits(x) == bits(1.0)) return 0.0;
    const double r = x - 1.0;
    const double r2 = r * r;
    const double r3 = r * r2;
    const double inner = __builtin_fma(r3, kB10, __builtin_fma(r2, kB9, __builtin_fma(r, kB8, kB7)));
    const double mid = __builtin_fma(r3, inner, __builtin_fma(r2, kB6, __builtin_fma(r, kB5, kB4)));
    const double poly = __builtin_fma(r3, mid, __builtin_fma(r2, kB3, __builtin_fma(r, kB2, kB1)));
    // r^2/2 exactly: r splits into a 26-bit head and its remainder
    double w = r * 0x1p27;
    const double rhi = r + w - w;
    const double rlo = r - rhi;
    w = rhi * rhi * kB0;
    const double hi = r + w;
    double lo = r - hi + w;
    lo = __builtin_fma(kB0 * rlo, rhi + r, lo);
    double y = __builtin_fma(r3, poly, lo);
    return y + hi;
}
}  // namespace logdetail

__host__ __device__ inline double libm_log(double x) {
    using namespace logdetail;
    uint64_t ix = bits(x);
    if (ix - bits(1.0 - 0x1p-4) < bits(1.0 + 0x1.09p-4) - bits(1.0 - 0x1p-4)) return near_one(x);
    const uint32_t top = (uint32_t)(ix >> 48);
    if (top - 0x0010u >= 0x7ff0u - 0x0010u) {
        if ((ix << 1) == 0) return -__builtin_inf();                  // log(+-0)
        if (ix == bits(__builtin_inf())) return x;                    // log(inf)
        if ((top & 0x8000u) || (top & 0x7ff0u) == 0x7ff0u) return __builtin_nan("");  // x < 0, nan
        ix = bits(x * 0x1p52) - (52ull << 52);                          // subnormal: normalise
    }
    const uint64_t tmp = ix - kOff;
    const int i = (int)((tmp >> 45) % 128);
    const int k = (int)((int64_t)tmp >> 52);
    const uint64_t iz = ix - (tmp & (0xfffull << 52));
    const double invc = from_bits(kLogTable[2 * i]);
    const double logc = from_bits(kLogTable[2 * i + 1]);
    const double z = from_bits(iz);
    const double r = __builtin_fma(z, invc, -1.0);
    const double kd = (double)k;
    const double w = __builtin_fma(kd, kLn2hi, logc);
    const double hi = w + r;
    const double lo = __builtin_fma(kd, kLn2lo, w - hi + r);
    const double r2 = r * r;
    const double p = __builtin_fma(r2, __builtin_fma(r, kA4, kA3), __builtin_fma(r, kA2, kA1));
    const double y = __builtin_fma(r * r2, p, __builtin_fma(r2, kA0, lo));
    return y + hi;
}

}  // namespace amcfd
