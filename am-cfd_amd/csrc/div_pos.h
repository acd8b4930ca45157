// Correctly rounded division on the dependency chain of the line sweeps'
// right-hand-side recurrence (CFDSolverManager.cpp:2442 and 2514, qr = (d + aw qr) /
// den), with the reciprocal y = RN(1/den) computed ahead of the sweep.
// tests/native/div_pos_check.cpp exercises it on the host.
#pragma once

#include <hip/hip_runtime.h>

namespace amcfd {

// x / b for b > 0, given y = RN(1/b), on the recurrence's dependency chain:
// a first quotient q0 = RN(x y), within 1.5 ulp of x/b, and one
// exact-remainder correction.  The exact value q0 - (b q0 - x) y differs
// from x/b by |1 - b y| |q0 - x/b| < 2^-53 * 1.5 ulp, so its rounding q1 is
// within an ulp of x/b, and is RN(x/b) unless x/b lies that close to a
// rounding midpoint.  The remainders are formed negated, b q - x, which for
// b > 0 leaves a zero x with the sign IEEE division gives it, so no select is
// needed.  Valid when b is positive and finite and x is zero or finite with
// |x| >= 2^-900 (the remainders are then exact).  Three dependent operations
// against the eleven of a full division; the caller confirms every row after
// the pass with div_pos_exact.
__host__ __device__ inline double div_pos(double x, double b, double y) {
    const double q0 = x * y;
    const double e0 = __builtin_fma(b, q0, -x);
    return __builtin_fma(-e0, y, q0);
}

// Whether q = div_pos(x, b, y) is RN(x / b).  Markstein's theorem (y the
// rounded reciprocal, q within an ulp of x/b) gives RN(q + (x - b q) y) =
// RN(x / b), so q is correctly rounded exactly when that second correction
// leaves it unchanged.  A zero x gives the zero of the right sign either way.
__host__ __device__ inline bool div_pos_exact(double x, double b, double y, double q) {
    const double e1 = __builtin_fma(b, q, -x);
    return __builtin_fma(-e1, y, q) == q;
}

}  // namespace amcfd
