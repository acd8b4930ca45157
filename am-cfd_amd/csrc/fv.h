// Device-side helpers shared by the HIP kernels: comparison semantics,
// convection-diffusion weighting, block reductions and the in-LDS
// tridiagonal solve.
#pragma once

#include "group.h"
#include "internal.h"
#include "libm_exp.h"
#include "libm_log.h"

namespace amcfd {

#define AMCFD_SMALL 1.0e-6
#define AMCFD_LARGE 1.0e20

// std::max / std::min semantics: the first argument wins ties and NaN
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }

// convection-diffusion weighting A(|P|), CFDSolverManager.cpp:3071-3101
__device__ __forceinline__ double weight_fn(int scheme, double pe) {
    if (scheme == AMCFD_DISC_POWERLAW) {
        double b = 1.0 - 0.1 * pe;
        double r = b * b;   // b^5 as ((((b*b)*b)*b)*b)
        r = r * b;
        r = r * b;
        r = r * b;
        return dmax(0.0, r);
    }
    if (scheme == AMCFD_DISC_EXPONENTIAL) return pe / (libm_exp(pe) - 1.0);
    return 1.0;
}
// neighbour coefficient D*A(|F|/D) + max(0, F) (upstream on the + side)
__device__ __forceinline__ double coef_plus(int scheme, double d, double f) {
    return d * weight_fn(scheme, fabs(f) / d) + dmax(0.0, f);
}
// neighbour coefficient D*A(|F|/D) + max(0, -F)
__device__ __forceinline__ double coef_minus(int scheme, double d, double f) {
    return d * weight_fn(scheme, fabs(f) / d) + dmax(0.0, -f);
}
// distance-weighted harmonic mean p*q / (wq*q + wp*p)
__device__ __forceinline__ double hmean(double p, double q, double wq, double wp) {
    return p * q / (wq * q + wp * p);
}

// wave64 sum / max via cross-lane shuffles (fixed tree: deterministic)
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off, 64));
    return v;
}

// block-wide sum of NV values per thread; result valid in thread 0.
// scratch must hold NV * (blockDim/64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *scratch) {
    const int lane = threadIdx.x + threadIdx.y * blockDim.x;
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if ((lane & 63) == 0)
#pragma unroll
        for (int q = 0; q < NV; ++q) scratch[q * nw + (lane >> 6)] = v[q];
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += scratch[q * nw + w];
            v[q] = s;
        }
    }
    __syncthreads();
}

#define HIP_CHECK_LAUNCH() (void)hipGetLastError()

}  // namespace amcfd
