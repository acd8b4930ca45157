// Device-side helpers shared by the HIP kernels: comparison semantics,
// convection-diffusion weighting, block reductions and the in-LDS
// tridiagonal solve.
#pragma once

#include "internal.h"
#include "libm_exp.h"

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

// Parallel cyclic reduction on an n-row tridiagonal system held in LDS:
//   a[i] x[i-1] + b[i] x[i] + c[i] x[i+1] = d[i],  a[0] = c[n-1] = 0.
// The solution overwrites d.  Every thread of the block must call it.
// Each thread owns rows tid, tid+nt, ... (at most MAXR rows).
template <int MAXR>
__device__ void pcr_solve(int n, double *a, double *b, double *c, double *d) {
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    for (int s = 1; s < n; s <<= 1) {
        double na[MAXR], nb[MAXR], nc[MAXR], nd[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            int i = tid + r * nt;
            if (i < n) {
                double ai = a[i], bi = b[i], ci = c[i], di = d[i];
                double k1 = 0.0, k2 = 0.0;
                double am = 0.0, cm = 0.0, dm = 0.0, ap = 0.0, cp = 0.0, dp = 0.0;
                if (i - s >= 0) { k1 = ai / b[i - s]; am = a[i - s]; cm = c[i - s]; dm = d[i - s]; }
                if (i + s < n) { k2 = ci / b[i + s]; ap = a[i + s]; cp = c[i + s]; dp = d[i + s]; }
                na[r] = -am * k1;
                nb[r] = bi - cm * k1 - ap * k2;
                nc[r] = -cp * k2;
                nd[r] = di - dm * k1 - dp * k2;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            int i = tid + r * nt;
            if (i < n) { a[i] = na[r]; b[i] = nb[r]; c[i] = nc[r]; d[i] = nd[r]; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
        int i = tid + r * nt;
        if (i < n) d[i] = d[i] / b[i];
    }
    __syncthreads();
}

#define HIP_CHECK_LAUNCH() (void)hipGetLastError()

}  // namespace amcfd
