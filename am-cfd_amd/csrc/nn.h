// Learned warm start: network kernels (nn.hip) and the host object that
// trains the network and applies it as a step's initial guess (warmstart.cpp).
#pragma once

#include "internal.h"

namespace amcfd {
namespace nn {

constexpr int NN_CIN = 16;    // input channels (8 features, zero padded to the MFMA width)
constexpr int NN_OUT = 2;     // predicted channels: temperature increment, pressure change
constexpr int NN_COUT = 16;   // head output channels (NN_OUT used)

struct ConvShape {
    int N, D, H, W, Cin;      // input
    int Do, Ho, Wo, Cout;     // output
    int stride;               // 1 or 2 (3x3x3 taps, padding 1)
};

// a box of cells of the solver's grid: origin (i0, j0, k0), extent W x H x D
struct Window {
    int i0, j0, k0, D, H, W;
};

// feature / target scaling
struct FeatScale {
    double t0, tinv, vinv, qinv;   // (T - t0) * tinv, velocity * vinv, flux * qinv
    double dtinv, dpinv;           // targets: dT * dtinv, dp * dpinv
    double dtscale, dpscale;       // inverse of dtinv, dpinv
};

struct GuessConst {
    double tliquid, tsolid, hlcal, hsmelt, acpa, acpb, acpl;
    double tfactor, pfactor;       // how much of the predicted increments is applied
};

struct AdamStep {
    float lr, b1, b2, eps, c1, c2;   // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
};

void conv_fwd(hipStream_t st, const ConvShape &s, const unsigned short *X, const unsigned short *W, const float *b,
              unsigned short *Y, float *Yf, int relu);
// dW += (FP32, atomics); dX += when dX is not null
void conv_bwd(hipStream_t st, const ConvShape &s, const unsigned short *X, const unsigned short *dZ,
              const unsigned short *WT, float *dW, float *dX);
void relu_grad(hipStream_t st, long long n, int C, const float *dA, const unsigned short *Y, int relu,
               unsigned short *dZ, float *db);
void up_cat(hipStream_t st, int N, int D, int H, int W, int C, int Cs, const unsigned short *lo,
            const unsigned short *skip, unsigned short *cat);
void up_cat_bwd(hipStream_t st, int N, int D, int H, int W, int C, int Cs, const float *dcat, float *dlo,
                float *dskip);
void mse(hipStream_t st, long long npos, int C, int nout, const float *out, const float *tgt, const float *mask,
         float inv_count, float *dout, float *acc);
void adam(hipStream_t st, int Cout, int Cin, float *w, float *m, float *v, const float *g, float *bw, float *bm,
          float *bv, const float *bg, const AdamStep &a, unsigned short *wb, unsigned short *wtb);
void weights_to_bf16(hipStream_t st, int Cout, int Cin, const float *w, unsigned short *wb, unsigned short *wtb);
void to_bf16(hipStream_t st, long long n, const float *a, unsigned short *b);
void features(hipStream_t st, const Dims &d, const Window &w, const FeatScale &f, const DevFields &fl,
              unsigned short *X, float *mask);
void target(hipStream_t st, const Dims &d, const Window &w, const FeatScale &f, const double *t_new,
            const double *t_old, const double *p_new, const double *p_old, float *tgt);
void apply_guess(hipStream_t st, const Dims &d, const Window &w, const FeatScale &f, const GuessConst &g,
                 const float *out, const DevFields &fl);

}  // namespace nn
}  // namespace amcfd
