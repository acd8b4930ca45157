// Kernels of the learned warm start: a 3-D convolutional encoder-decoder on
// BF16 matrix cores (gfx950 MFMA, FP32 accumulation) with its training pass.
//
// Tensors are NDHWC (channels fastest), BF16 stored as 16-bit words, every
// channel count a multiple of 16.  A 3x3x3 convolution with padding 1 and
// stride s is an implicit GEMM: rows = output positions, columns = output
// channels, reduction = 27 taps x input channels.  One wave computes a 16 x
// (16 * NT) tile with v_mfma_f32_16x16x16_bf16: lane l feeds row (l % 16) of
// A and column (l % 16) of B with the four reduction indices 4 (l / 16) ..
// 4 (l / 16) + 3, and receives rows 4 (l / 16) .. +3 of column l % 16.
// Channels-last makes every A and B fragment one 8-byte load.
//   forward        X[pos][ci] (*) W[co][tap][ci]       -> Y[pos][co] (+bias, ReLU)
//   backward data  dZ[pos][co] (*) WT[ci][tap][co]     -> dX[pos][ci] (accumulated)
//   backward weight sum over positions dZ[.][co] X[.][ci] -> dW[co][tap][ci] (FP32 atomics)

#include "nn.h"

namespace amcfd {
namespace nn {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);   // inf / nan kept
    u += 0x7fffu + ((u >> 16) & 1u);                                          // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// output position p (flattened n, z, y, x) -> coordinates
struct Pos {
    int n, z, y, x;
};
__device__ __forceinline__ Pos unflat(long long p, int D, int H, int W) {
    Pos q;
    q.x = (int)(p % W);
    p /= W;
    q.y = (int)(p % H);
    p /= H;
    q.z = (int)(p % D);
    q.n = (int)(p / D);
    return q;
}

template <int NT>
__global__ __launch_bounds__(64) void k_conv_fwd(ConvShape s, const unsigned short *__restrict__ X,
                                                  const unsigned short *__restrict__ Wt,
                                                  const float *__restrict__ bias, unsigned short *__restrict__ Y,
                                                  float *__restrict__ Yf, int relu) {
    const int lane = threadIdx.x, i = lane & 15, kq = lane >> 4;
    const long long npos = (long long)s.N * s.Do * s.Ho * s.Wo;
    const long long p0 = (long long)blockIdx.x * 16;
    const int n0 = blockIdx.y * 16 * NT;
    const long long prow = p0 + i;
    const Pos o = unflat(prow < npos ? prow : npos - 1, s.Do, s.Ho, s.Wo);
    v4f acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 27; ++tap) {
        const int tz = tap / 9, ty = (tap / 3) % 3, tx = tap % 3;
        const int zi = o.z * s.stride + tz - 1, yi = o.y * s.stride + ty - 1, xi = o.x * s.stride + tx - 1;
        const bool ok = prow < npos && zi >= 0 && zi < s.D && yi >= 0 && yi < s.H && xi >= 0 && xi < s.W;
        const size_t xin = ((((size_t)o.n * s.D + (ok ? zi : 0)) * s.H + (ok ? yi : 0)) * s.W + (ok ? xi : 0)) * s.Cin;
        for (int c0 = 0; c0 < s.Cin; c0 += 16) {
            v4s a = {0, 0, 0, 0};
            if (ok) a = *(const v4s *)(X + xin + c0 + 4 * kq);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int co = n0 + 16 * t + i;
                const v4s b = *(const v4s *)(Wt + ((size_t)co * 27 + tap) * s.Cin + c0 + 4 * kq);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = n0 + 16 * t + i;   // column of this lane
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long long p = p0 + 4 * kq + e;   // row of this lane's e-th value
            if (p >= npos) continue;
            float v = acc[t][e] + bias[co];
            if (relu) v = fmaxf(v, 0.f);
            if (Y) Y[(size_t)p * s.Cout + co] = f2bf(v);
            if (Yf) Yf[(size_t)p * s.Cout + co] = v;
        }
    }
}

// dX[pos_in][ci] += sum over taps and co of WT[ci][tap][co] dZ[pos_out][co],
// pos_out the output position that reads pos_in through the tap
template <int NT>
__global__ __launch_bounds__(64) void k_conv_bwd_data(ConvShape s, const unsigned short *__restrict__ dZ,
                                                       const unsigned short *__restrict__ WT, float *__restrict__ dX) {
    const int lane = threadIdx.x, i = lane & 15, kq = lane >> 4;
    const long long npos = (long long)s.N * s.D * s.H * s.W;
    const long long p0 = (long long)blockIdx.x * 16;
    const int n0 = blockIdx.y * 16 * NT;
    const long long prow = p0 + i;
    const Pos q = unflat(prow < npos ? prow : npos - 1, s.D, s.H, s.W);
    v4f acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 27; ++tap) {
        const int tz = tap / 9, ty = (tap / 3) % 3, tx = tap % 3;
        // zo * stride + tz - 1 = zi
        const int az = q.z + 1 - tz, ay = q.y + 1 - ty, ax = q.x + 1 - tx;
        bool ok = prow < npos && az >= 0 && ay >= 0 && ax >= 0;
        if (s.stride == 2) ok = ok && !(az & 1) && !(ay & 1) && !(ax & 1);
        const int zo = s.stride == 2 ? az >> 1 : az, yo = s.stride == 2 ? ay >> 1 : ay, xo = s.stride == 2 ? ax >> 1 : ax;
        ok = ok && zo < s.Do && yo < s.Ho && xo < s.Wo;
        const size_t zin = ((((size_t)q.n * s.Do + (ok ? zo : 0)) * s.Ho + (ok ? yo : 0)) * s.Wo + (ok ? xo : 0)) * s.Cout;
        for (int c0 = 0; c0 < s.Cout; c0 += 16) {
            v4s a = {0, 0, 0, 0};
            if (ok) a = *(const v4s *)(dZ + zin + c0 + 4 * kq);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int ci = n0 + 16 * t + i;
                const v4s b = *(const v4s *)(WT + ((size_t)ci * 27 + tap) * s.Cout + c0 + 4 * kq);
                acc[t] = mfma(a, b, acc[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int ci = n0 + 16 * t + i;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long long p = p0 + 4 * kq + e;
            if (p < npos) dX[(size_t)p * s.Cin + ci] += acc[t][e];
        }
    }
}

// dW[co][tap][ci] += sum over a chunk of output positions of dZ[p][co] *
// X[in(p, tap)][ci]: rows co, columns ci, reduction over positions
constexpr int BW_CHUNK = 256;

__global__ __launch_bounds__(64) void k_conv_bwd_weight(ConvShape s, const unsigned short *__restrict__ dZ,
                                                         const unsigned short *__restrict__ X, float *__restrict__ dW) {
    const int lane = threadIdx.x, i = lane & 15, kq = lane >> 4;
    const long long npos = (long long)s.N * s.Do * s.Ho * s.Wo;
    const int nco = s.Cout / 16;
    int t = blockIdx.y;
    const int tap = t % 27;
    t /= 27;
    const int cot = t % nco, cit = t / nco;
    const int tz = tap / 9, ty = (tap / 3) % 3, tx = tap % 3;
    const int co = cot * 16 + i, ci = cit * 16 + i;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    const long long pb = (long long)blockIdx.x * BW_CHUNK, pe = min(pb + BW_CHUNK, npos);
    for (long long p0 = pb; p0 < pe; p0 += 16) {
        short av[4], bv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long long p = p0 + 4 * kq + e;
            av[e] = 0;
            bv[e] = 0;
            if (p < pe) {
                const Pos o = unflat(p, s.Do, s.Ho, s.Wo);
                av[e] = (short)dZ[(size_t)p * s.Cout + co];
                const int zi = o.z * s.stride + tz - 1, yi = o.y * s.stride + ty - 1, xi = o.x * s.stride + tx - 1;
                if (zi >= 0 && zi < s.D && yi >= 0 && yi < s.H && xi >= 0 && xi < s.W)
                    bv[e] = (short)X[((((size_t)o.n * s.D + zi) * s.H + yi) * s.W + xi) * s.Cin + ci];
            }
        }
        const v4s a = {av[0], av[1], av[2], av[3]}, b = {bv[0], bv[1], bv[2], bv[3]};
        acc = mfma(a, b, acc);
    }
    // acc[e]: row co = cot*16 + 4 kq + e, column ci = cit*16 + i
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int r = cot * 16 + 4 * kq + e;
        atomicAdd(&dW[((size_t)r * 27 + tap) * s.Cin + ci], acc[e]);
    }
}

// dZ = dA * relu'(Y) (or dA alone), BF16; db += sum over positions
__global__ __launch_bounds__(256) void k_relu_grad(long long n, int C, const float *__restrict__ dA,
                                                   const unsigned short *__restrict__ Y, int relu,
                                                   unsigned short *__restrict__ dZ, float *__restrict__ db) {
    __shared__ float sdb[256];
    const int c = threadIdx.x % C;   // blockDim is a multiple of C
    float acc = 0.f;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        float g = dA[e];
        if (relu && !(bf2f(Y[e]) > 0.f)) g = 0.f;
        const unsigned short gb = f2bf(g);
        dZ[e] = gb;
        acc += bf2f(gb);
    }
    sdb[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < C) {
        float s = 0.f;
        for (int t = threadIdx.x; t < (int)blockDim.x; t += C) s += sdb[t];
        atomicAdd(&db[c], s);
    }
}

// nearest x2 upsampling of lo (C channels) into channels [0, C) of cat
// (Ccat channels), and a copy of skip (Cs channels) into [C, C + Cs)
__global__ __launch_bounds__(256) void k_up_cat(int N, int D, int H, int W, int C, int Cs,
                                                const unsigned short *__restrict__ lo,
                                                const unsigned short *__restrict__ skip, unsigned short *__restrict__ cat) {
    const int Cc = C + Cs;
    const long long n = (long long)N * D * H * W * Cc;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
        const int c = (int)(e % Cc);
        const long long p = e / Cc;
        if (c < C) {
            const Pos q = unflat(p, D, H, W);
            const size_t lp = (((size_t)q.n * (D / 2) + q.z / 2) * (H / 2) + q.y / 2) * (W / 2) + q.x / 2;
            cat[e] = lo[lp * C + c];
        } else {
            cat[e] = skip[(size_t)p * Cs + (c - C)];
        }
    }
}

// backward of k_up_cat: dlo += sum over the 8 children, dskip += its channels
__global__ __launch_bounds__(256) void k_up_cat_bwd(int N, int D, int H, int W, int C, int Cs,
                                                    const float *__restrict__ dcat, float *__restrict__ dlo,
                                                    float *__restrict__ dskip) {
    const int Cc = C + Cs;
    const long long nlo = (long long)N * (D / 2) * (H / 2) * (W / 2) * C;
    const long long nsk = (long long)N * D * H * W * Cs;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nlo + nsk; e += (long long)gridDim.x * 256) {
        if (e < nlo) {
            const int c = (int)(e % C);
            const Pos q = unflat(e / C, D / 2, H / 2, W / 2);
            float s = 0.f;
            for (int dz = 0; dz < 2; ++dz)
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        const size_t p = (((size_t)q.n * D + 2 * q.z + dz) * H + 2 * q.y + dy) * W + 2 * q.x + dx;
                        s += dcat[p * Cc + c];
                    }
            dlo[e] += s;
        } else {
            const long long f = e - nlo;
            const int c = (int)(f % Cs);
            dskip[f] += dcat[(size_t)(f / Cs) * Cc + C + c];
        }
    }
}

// masked mean-squared error over the first NOUT channels of the FP32 head
// output; writes dOut (FP32 over all C channels, zero beyond NOUT) and adds
// the masked sum of squares to acc
__global__ __launch_bounds__(256) void k_mse(long long npos, int C, int nout, const float *__restrict__ out,
                                             const float *__restrict__ tgt, const float *__restrict__ mask,
                                             float inv_count, float *__restrict__ dout, float *__restrict__ acc) {
    __shared__ float sl[256];
    float l = 0.f;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < npos; p += (long long)gridDim.x * 256) {
        const float m = mask[p];
        for (int c = 0; c < C; ++c) {
            float g = 0.f;
            if (c < nout) {
                const float d = out[(size_t)p * C + c] - tgt[(size_t)p * nout + c];
                l += m * d * d;
                g = 2.f * m * d * inv_count;
            }
            dout[(size_t)p * C + c] = g;
        }
    }
    sl[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sl[threadIdx.x] += sl[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc, sl[0]);
}

// Adam on the FP32 master weights; refreshes the BF16 copies W[co][tap][ci]
// and WT[ci][tap][co]
__global__ __launch_bounds__(256) void k_adam(int Cout, int Cin, float *__restrict__ w, float *__restrict__ m,
                                              float *__restrict__ v, const float *__restrict__ g, float lr, float b1,
                                              float b2, float eps, float c1, float c2, unsigned short *__restrict__ wb,
                                              unsigned short *__restrict__ wtb) {
    const long long n = (long long)Cout * 27 * Cin;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
        const float gg = g[e];
        const float mm = b1 * m[e] + (1.f - b1) * gg, vv = b2 * v[e] + (1.f - b2) * gg * gg;
        m[e] = mm;
        v[e] = vv;
        const float nw = w[e] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
        w[e] = nw;
        const int ci = (int)(e % Cin), tap = (int)((e / Cin) % 27), co = (int)(e / ((long long)Cin * 27));
        const unsigned short b = f2bf(nw);
        wb[e] = b;
        wtb[((size_t)ci * 27 + tap) * Cout + co] = b;
    }
}

__global__ __launch_bounds__(256) void k_adam_bias(int n, float *__restrict__ w, float *__restrict__ m,
                                                   float *__restrict__ v, const float *__restrict__ g, float lr, float b1,
                                                   float b2, float eps, float c1, float c2) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg, vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    w[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

__global__ __launch_bounds__(256) void k_to_bf16(long long n, const float *__restrict__ a, unsigned short *__restrict__ b) {
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) b[e] = f2bf(a[e]);
}

__global__ __launch_bounds__(256) void k_transpose_w(int Cout, int Cin, const unsigned short *__restrict__ wb,
                                                     unsigned short *__restrict__ wtb) {
    const long long n = (long long)Cout * 27 * Cin;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
        const int ci = (int)(e % Cin), tap = (int)((e / Cin) % 27), co = (int)(e / ((long long)Cin * 27));
        wtb[((size_t)ci * 27 + tap) * Cout + co] = wb[e];
    }
}

// ---------------------------------------------------------------------------
// solver state <-> network tensors.  A window of WD x WH x WW cells (z, y, x)
// with origin (k0, j0, i0); cells outside the allocated field read as zero
// and are masked.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_features(Dims d, Window w, FeatScale f, const double *__restrict__ temp,
                                                  const double *__restrict__ u, const double *__restrict__ v,
                                                  const double *__restrict__ wv, const double *__restrict__ fl,
                                                  const double *__restrict__ heatin, unsigned short *__restrict__ X,
                                                  float *__restrict__ mask) {
    const long long n = (long long)w.D * w.H * w.W;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
        const int x = (int)(p % w.W), y = (int)((p / w.W) % w.H), z = (int)(p / ((long long)w.W * w.H));
        const int i = w.i0 + x, j = w.j0 + y, k = w.k0 + z;
        float ft[NN_CIN] = {0};
        const bool in = i >= 0 && i < d.ni && j >= 0 && j < d.nj && k >= 0 && k < d.nk;
        const bool interior = i >= 1 && i <= d.nim2 && j >= 1 && j <= d.njm2 && k >= 1 && k <= d.nkm2;
        if (in) {
            const size_t c = (size_t)k * d.sk + (size_t)j * d.sj + i;
            ft[0] = (float)((temp[c] - f.t0) * f.tinv);
            ft[1] = (float)(u[c] * f.vinv);
            ft[2] = (float)(v[c] * f.vinv);
            ft[3] = (float)(wv[c] * f.vinv);
            ft[4] = (float)fl[c];
            ft[5] = (float)(heatin[(size_t)j * d.ni + i] * f.qinv);
            ft[6] = 1.f;   // inside the field
            ft[7] = (float)(d.nkm1 - k) * (1.f / 16.f);   // depth below the top
        }
        for (int c = 0; c < NN_CIN; ++c) X[(size_t)p * NN_CIN + c] = f2bf(ft[c]);
        if (mask) mask[p] = interior ? 1.f : 0.f;
    }
}

// training target: (T_new - T_old) / dT_scale and (p_new - p_old) / p_scale
__global__ __launch_bounds__(256) void k_target(Dims d, Window w, FeatScale f, const double *__restrict__ t_new,
                                                const double *__restrict__ t_old, const double *__restrict__ p_new,
                                                const double *__restrict__ p_old, float *__restrict__ tgt) {
    const long long n = (long long)w.D * w.H * w.W;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
        const int x = (int)(p % w.W), y = (int)((p / w.W) % w.H), z = (int)(p / ((long long)w.W * w.H));
        const int i = w.i0 + x, j = w.j0 + y, k = w.k0 + z;
        float a = 0.f, b = 0.f;
        if (i >= 0 && i < d.ni && j >= 0 && j < d.nj && k >= 0 && k < d.nk) {
            const size_t c = (size_t)k * d.sk + (size_t)j * d.sj + i;
            a = (float)((t_new[c] - t_old[c]) * f.dtinv);
            b = (float)((p_new[c] - p_old[c]) * f.dpinv);
        }
        tgt[p * NN_OUT] = a;
        tgt[p * NN_OUT + 1] = b;
    }
}

// initial guess of the step: T + predicted increment mapped back to
// enthalpy through the inverse of convertEnthalpyToTemp
// (CFDSolverManager.cpp:2550-2570), and p + the predicted pressure change
__global__ __launch_bounds__(256) void k_apply_guess(Dims d, Window w, FeatScale f, GuessConst g,
                                                     const float *__restrict__ out, const double *__restrict__ temp,
                                                     double *__restrict__ h, double *__restrict__ pres) {
    const long long n = (long long)w.D * w.H * w.W;
    for (long long p = (long long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
        const int x = (int)(p % w.W), y = (int)((p / w.W) % w.H), z = (int)(p / ((long long)w.W * w.H));
        const int i = w.i0 + x, j = w.j0 + y, k = w.k0 + z;
        if (!(i >= 1 && i <= d.nim2 && j >= 1 && j <= d.njm2 && k >= 1 && k <= d.nkm2)) continue;
        const size_t c = (size_t)k * d.sk + (size_t)j * d.sj + i;
        const double dT = (double)out[(size_t)p * NN_COUT] * f.dtscale * g.tfactor;
        const double T = temp[c] + dT;
        double hv;
        if (T >= g.tliquid) hv = g.hlcal + g.acpl * (T - g.tliquid);
        else if (T <= g.tsolid) hv = 0.5 * g.acpa * T * T + g.acpb * T;
        else hv = g.hsmelt + (T - g.tsolid) / (g.tliquid - g.tsolid) * (g.hlcal - g.hsmelt);
        h[c] = hv;
        if (g.pfactor != 0.0) pres[c] += (double)out[(size_t)p * NN_COUT + 1] * f.dpscale * g.pfactor;
    }
}

template <int NT>
void conv_fwd_t(hipStream_t st, const ConvShape &s, const unsigned short *X, const unsigned short *W, const float *b,
                unsigned short *Y, float *Yf, int relu) {
    const long long npos = (long long)s.N * s.Do * s.Ho * s.Wo;
    hipLaunchKernelGGL(k_conv_fwd<NT>, dim3((unsigned)((npos + 15) / 16), s.Cout / (16 * NT)), dim3(64), 0, st, s, X,
                       W, b, Y, Yf, relu);
}
template <int NT>
void conv_bwd_data_t(hipStream_t st, const ConvShape &s, const unsigned short *dZ, const unsigned short *WT, float *dX) {
    const long long npos = (long long)s.N * s.D * s.H * s.W;
    hipLaunchKernelGGL(k_conv_bwd_data<NT>, dim3((unsigned)((npos + 15) / 16), s.Cin / (16 * NT)), dim3(64), 0, st, s,
                       dZ, WT, dX);
}

unsigned blocks_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 8192); }

}  // namespace

void conv_fwd(hipStream_t st, const ConvShape &s, const unsigned short *X, const unsigned short *W, const float *b,
              unsigned short *Y, float *Yf, int relu) {
    if (s.Cout % 64 == 0) conv_fwd_t<4>(st, s, X, W, b, Y, Yf, relu);
    else if (s.Cout % 32 == 0) conv_fwd_t<2>(st, s, X, W, b, Y, Yf, relu);
    else conv_fwd_t<1>(st, s, X, W, b, Y, Yf, relu);
}

void conv_bwd(hipStream_t st, const ConvShape &s, const unsigned short *X, const unsigned short *dZ,
              const unsigned short *WT, float *dW, float *dX) {
    const long long nout = (long long)s.N * s.Do * s.Ho * s.Wo;
    const unsigned nch = (unsigned)((nout + BW_CHUNK - 1) / BW_CHUNK);
    hipLaunchKernelGGL(k_conv_bwd_weight, dim3(nch, 27 * (s.Cout / 16) * (s.Cin / 16)), dim3(64), 0, st, s, dZ, X, dW);
    if (dX) {
        if (s.Cin % 64 == 0) conv_bwd_data_t<4>(st, s, dZ, WT, dX);
        else if (s.Cin % 32 == 0) conv_bwd_data_t<2>(st, s, dZ, WT, dX);
        else conv_bwd_data_t<1>(st, s, dZ, WT, dX);
    }
}

void relu_grad(hipStream_t st, long long n, int C, const float *dA, const unsigned short *Y, int relu,
               unsigned short *dZ, float *db) {
    const int bs = (256 / C) * C;
    hipLaunchKernelGGL(k_relu_grad, dim3(std::min<long long>((n + bs - 1) / bs, 4096)), dim3(bs), 0, st, n, C, dA, Y,
                       relu, dZ, db);
}

void up_cat(hipStream_t st, int N, int D, int H, int W, int C, int Cs, const unsigned short *lo,
            const unsigned short *skip, unsigned short *cat) {
    hipLaunchKernelGGL(k_up_cat, dim3(blocks_for((long long)N * D * H * W * (C + Cs))), dim3(256), 0, st, N, D, H, W, C,
                       Cs, lo, skip, cat);
}

void up_cat_bwd(hipStream_t st, int N, int D, int H, int W, int C, int Cs, const float *dcat, float *dlo,
                float *dskip) {
    const long long n = (long long)N * (D / 2) * (H / 2) * (W / 2) * C + (long long)N * D * H * W * Cs;
    hipLaunchKernelGGL(k_up_cat_bwd, dim3(blocks_for(n)), dim3(256), 0, st, N, D, H, W, C, Cs, dcat, dlo, dskip);
}

void mse(hipStream_t st, long long npos, int C, int nout, const float *out, const float *tgt, const float *mask,
         float inv_count, float *dout, float *acc) {
    hipLaunchKernelGGL(k_mse, dim3(blocks_for(npos)), dim3(256), 0, st, npos, C, nout, out, tgt, mask, inv_count, dout,
                       acc);
}

void adam(hipStream_t st, int Cout, int Cin, float *w, float *m, float *v, const float *g, float *bw, float *bm,
          float *bv, const float *bg, const AdamStep &a, unsigned short *wb, unsigned short *wtb) {
    const long long n = (long long)Cout * 27 * Cin;
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, st, Cout, Cin, w, m, v, g, a.lr, a.b1, a.b2, a.eps,
                       a.c1, a.c2, wb, wtb);
    hipLaunchKernelGGL(k_adam_bias, dim3((Cout + 255) / 256), dim3(256), 0, st, Cout, bw, bm, bv, bg, a.lr, a.b1, a.b2,
                       a.eps, a.c1, a.c2);
}

void weights_to_bf16(hipStream_t st, int Cout, int Cin, const float *w, unsigned short *wb, unsigned short *wtb) {
    const long long n = (long long)Cout * 27 * Cin;
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, st, n, w, wb);
    hipLaunchKernelGGL(k_transpose_w, dim3(blocks_for(n)), dim3(256), 0, st, Cout, Cin, wb, wtb);
}

void to_bf16(hipStream_t st, long long n, const float *a, unsigned short *b) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, st, n, a, b);
}

void features(hipStream_t st, const Dims &d, const Window &w, const FeatScale &f, const DevFields &fl,
              unsigned short *X, float *mask) {
    hipLaunchKernelGGL(k_features, dim3(blocks_for((long long)w.D * w.H * w.W)), dim3(256), 0, st, d, w, f, fl.temp,
                       fl.uvel, fl.vvel, fl.wvel, fl.fracl, fl.heatin, X, mask);
}

void target(hipStream_t st, const Dims &d, const Window &w, const FeatScale &f, const double *t_new,
            const double *t_old, const double *p_new, const double *p_old, float *tgt) {
    hipLaunchKernelGGL(k_target, dim3(blocks_for((long long)w.D * w.H * w.W)), dim3(256), 0, st, d, w, f, t_new, t_old,
                       p_new, p_old, tgt);
}

void apply_guess(hipStream_t st, const Dims &d, const Window &w, const FeatScale &f, const GuessConst &g,
                 const float *out, const DevFields &fl) {
    hipLaunchKernelGGL(k_apply_guess, dim3(blocks_for((long long)w.D * w.H * w.W)), dim3(256), 0, st, d, w, f, g, out,
                       fl.temp, fl.enthalpy, fl.pressure);
}

}  // namespace nn
}  // namespace amcfd
