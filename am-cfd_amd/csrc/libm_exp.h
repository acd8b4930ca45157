// exp() reproducing the C library's result bit for bit, for host and device.
//
// The reference calls std::exp (BoundCondManager.cpp:150, :604, :729,
// CFDSolverManager.cpp:1735, :3100), which on its x86-64 Linux hosts is
// glibc 2.35's exp in the FMA ifunc variant.  That routine is the published
// table-driven algorithm of Szabolcs Nagy (ARM optimized-routines, 2018):
//   x = k ln2/128 + r,  |r| <= ln2/256,
//   exp(x) = 2^(k/128) * exp(r) ~= scale + scale * (tail + r + r^2 P(r)),
// with a 128-entry table of 2^(j/128) (see gen_exp_table.py), a degree-5
// polynomial and a separate scaling path for 512 <= |x| < 1024.  It is
// accurate to about 0.51 ulp and is not correctly rounded (about 1 result in
// 1300 differs from the correctly rounded one), so a device exp of any other
// construction disagrees with the host in the last bit often enough to
// perturb a SIMPLE run.  This is the same algorithm with the same fused
// multiply-adds, written for both sides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "exp_table.h"

namespace amcfd {

namespace expdetail {
__host__ __device__ inline uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
__host__ __device__ inline double from_bits(uint64_t u) { return __builtin_bit_cast(double, u); }
__host__ __device__ inline uint32_t top12(double x) { return (uint32_t)(bits(x) >> 52); }

constexpr double kInvLn2N = 0x1.71547652b82fep0 * 128;  // 128 / ln 2
constexpr double kNegLn2hiN = -0x1.62e42fefa0000p-8;    // -ln2/128, high part (exact products)
constexpr double kNegLn2loN = -0x1.cf79abc9e3b3ap-47;   // -ln2/128, low part
constexpr double kShift = 0x1.8p52;                      // rounds k ln2/128 to an integer
// minimax coefficients of (exp(r) - 1 - r) / r^2 on |r| <= ln2/256
constexpr double kC2 = 0x1.ffffffffffdbdp-2, kC3 = 0x1.555555555543cp-3;
constexpr double kC4 = 0x1.55555cf172b91p-5, kC5 = 0x1.1111167a4d017p-7;

// 512 <= |x| < 1024: scale may over- or underflow, so rescale first
__host__ __device__ inline double scaled(double tmp, uint64_t sbits, uint64_t ki) {
    if ((ki & 0x80000000ull) == 0) {
        const double scale = from_bits(sbits - (1009ull << 52));
        return 0x1p1009 * __builtin_fma(scale, tmp, scale);
    }
    // here the product scale * tmp is rounded once and used twice (the C
    // library's build fuses a product into an add only when it has one use)
    const double scale = from_bits(sbits + (1022ull << 52));
    const double st = scale * tmp;
    double y = scale + st;
    if (y < 1.0) {
        // round in the normal range before scaling into the subnormal range
        double lo = scale - y + st;
        const double hi = 1.0 + y;
        lo = 1.0 - hi + y + lo;
        y = (hi + lo) - 1.0;
        if (y == 0.0) y = 0.0;
    }
    return 0x1p-1022 * y;
}
}  // namespace expdetail

__host__ __device__ inline double libm_exp(double x) {
    using namespace expdetail;
    uint32_t abstop = top12(x) & 0x7ff;
    if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54)) {
        if ((int32_t)(abstop - top12(0x1p-54)) < 0) return 1.0 + x;   // |x| < 2^-54
        if (abstop >= top12(1024.0)) {
            if (bits(x) == bits(-__builtin_inf())) return 0.0;
            if (abstop >= top12(__builtin_inf())) return 1.0 + x;     // +inf, nan
            return (bits(x) >> 63) ? 0.0 : __builtin_inf();
        }
        abstop = 0;   // 512 <= |x| < 1024
    }
    double kd = __builtin_fma(kInvLn2N, x, kShift);
    const uint64_t ki = bits(kd);
    kd -= kShift;
    const double r = __builtin_fma(kd, kNegLn2loN, __builtin_fma(kd, kNegLn2hiN, x));
    const uint64_t idx = 2 * (ki % 128);
    const uint64_t top = ki << 45;
    const double tail = from_bits(kExpTable[idx]);
    const uint64_t sbits = kExpTable[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = __builtin_fma(r2 * r2, __builtin_fma(r, kC5, kC4),
                                     __builtin_fma(r2, __builtin_fma(r, kC3, kC2), tail + r));
    if (abstop == 0) return scaled(tmp, sbits, ki);
    const double scale = from_bits(sbits);
    return __builtin_fma(scale, tmp, scale);
}

}  // namespace amcfd
