// Learned warm start: a 3-D convolutional encoder-decoder (4 levels, 16 to
// 128 channels, 3x3x3 kernels, BF16 weights and activations, FP32
// accumulation on the matrix cores) that maps the state at the start of a
// step (T, u, v, w, liquid fraction, the step's laser flux) in a window
// around the beam to the step's temperature increment and pressure change.
// It is trained on the GPU from steps the solver itself runs (record mode)
// and then supplies each step's initial iterate (apply mode).  The SIMPLE
// iteration (CFDSolverManager.cpp:595-722) still runs to the same
// convergence test, so the method still decides the answer: the guess can only
// change where the iteration starts (CFDSolverManager.cpp:608-612).
//
// Network (input D x H x W, D, H, W multiples of 8):
//   L0 16->16 s1, L1 16->32 s2, L2 32->64 s2, L3 64->128 s2, L4 128->128 s1
//   L5 [up(L4) 128 + L2 64] -> 64, L6 [up(L5) 64 + L1 32] -> 32,
//   L7 [up(L6) 32 + L0 16] -> 16, L8 16 -> 16 (2 used, linear)
// ReLU after every layer but L8; nearest x2 upsampling; Adam on FP32
// master weights.

#include "nn.h"
#include "host.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <unordered_map>

namespace amcfd {

#define WSCALL(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return set_error(AMCFD_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct LayerDef {
    int cin, cout, stride, relu;
};
constexpr int NL = 9;
constexpr LayerDef DEFS[NL] = {{16, 16, 1, 1},  {16, 32, 2, 1},  {32, 64, 2, 1}, {64, 128, 2, 1}, {128, 128, 1, 1},
                               {192, 64, 1, 1}, {96, 32, 1, 1},  {48, 16, 1, 1}, {16, 16, 1, 0}};
// sample window (z, y, x)
constexpr int WD = 16, WH = 32, WW = 32;

}  // namespace

class WarmStart {
  public:
    ~WarmStart();
    int create(int seed);
    int param_count() const;
    int get_params(float *host, size_t n);
    int set_params(const float *host, size_t n);
    int save(const char *path);
    int load(const char *path);
    // forward + loss + backward on host tensors (no update): test entry
    int loss_grad(const float *x, const float *tgt, const float *mask, int N, int D, int H, int W, double *loss,
                  float *grad);
    int train(int iters, int batch, double lr, double *loss_first, double *loss_last);
    int samples() {
        std::lock_guard<std::mutex> lock(mu_);
        return nsamp_;
    }
    // solver hooks (on the solver's stream); solvers on other host threads
    // may record and apply at the same time
    int capture_input(Solver &s);
    int capture_target(Solver &s);
    int apply(Solver &s);
    void attach(Solver *s);
    void detach(Solver *s);
    double tfactor = 1.0, pfactor = 0.0;

  private:
    struct Layer {
        LayerDef def;
        float *w = nullptr, *b = nullptr, *m = nullptr, *v = nullptr, *bm = nullptr, *bv = nullptr;
        float *gw = nullptr, *gb = nullptr;
        unsigned short *wb = nullptr, *wtb = nullptr;
        size_t nw() const { return (size_t)def.cout * 27 * def.cin; }
    } L_[NL];
    // activations of the current batch shape
    int N_ = 0, D_ = 0, H_ = 0, W_ = 0;
    unsigned short *x_ = nullptr, *a_[8] = {}, *c_[3] = {}, *dz_ = nullptr;
    float *out_ = nullptr, *dout_ = nullptr, *da_[8] = {}, *dc_[3] = {}, *tgt_ = nullptr, *mask_ = nullptr;
    float *acc_ = nullptr;
    // recorded samples (window-sized ring of cap_ slots; filled_[slot] once
    // a sample's target is written)
    int cap_ = 0, nsamp_ = 0, next_ = 0;
    unsigned short *sx_ = nullptr;
    float *stgt_ = nullptr, *smask_ = nullptr;
    std::vector<char> filled_;
    // per attached solver: the pending sample's start-of-step T and p (full
    // fields), its window and its slot (-1: none pending)
    struct Rec {
        double *told = nullptr, *pold = nullptr;
        size_t n = 0;
        nn::Window win{};
        int slot = -1;
    };
    std::unordered_map<Solver *, Rec> recs_;
    // guards the ring, recs_ and the network's buffers (one forward/train at
    // a time)
    std::mutex mu_;
    int step_t_ = 0;
    hipStream_t st_ = nullptr;
    std::vector<void *> allocs_, act_allocs_;
    std::mt19937 rng_;

    template <typename T>
    int dalloc(T **p, size_t n, std::vector<void *> &list) {
        void *v = nullptr;
        WSCALL(hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T)));
        WSCALL(hipMemsetAsync(v, 0, std::max<size_t>(n, 1) * sizeof(T), st_));
        list.push_back(v);
        *p = (T *)v;
        return 0;
    }
    int ensure_batch(int N, int D, int H, int W);
    nn::ConvShape shape(int l) const;
    void forward(hipStream_t st);
    int backward(hipStream_t st, float inv_count);
    nn::FeatScale scale(const Solver &s) const;
    nn::Window window(const Solver &s) const;
};

WarmStart::~WarmStart() {
    // solvers still attached forget this network (their recording buffers
    // are released with it)
    for (auto &kv : recs_) {
        (void)hipStreamSynchronize(kv.first->stream());
        kv.first->forget_warm_start();
        (void)hipFree(kv.second.told);
        (void)hipFree(kv.second.pold);
    }
    if (st_) (void)hipStreamSynchronize(st_);
    for (void *p : act_allocs_) (void)hipFree(p);
    for (void *p : allocs_) (void)hipFree(p);
    if (st_) (void)hipStreamDestroy(st_);
}

int WarmStart::create(int seed) {
    WSCALL(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    rng_.seed((unsigned)seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int l = 0; l < NL; ++l) {
        Layer &y = L_[l];
        y.def = DEFS[l];
        const size_t nw = y.nw(), nb = y.def.cout;
        int st;
        if ((st = dalloc(&y.w, nw, allocs_)) || (st = dalloc(&y.m, nw, allocs_)) || (st = dalloc(&y.v, nw, allocs_)) ||
            (st = dalloc(&y.gw, nw, allocs_)) || (st = dalloc(&y.b, nb, allocs_)) || (st = dalloc(&y.bm, nb, allocs_)) ||
            (st = dalloc(&y.bv, nb, allocs_)) || (st = dalloc(&y.gb, nb, allocs_)) || (st = dalloc(&y.wb, nw, allocs_)) ||
            (st = dalloc(&y.wtb, nw, allocs_)))
            return st;
        // He initialisation; the head starts small so the first guesses are
        // close to "no change"
        const float sd = std::sqrt(2.f / (27.f * y.def.cin)) * (y.def.relu ? 1.f : 0.1f);
        std::vector<float> w(nw);
        for (auto &v : w) v = sd * nd(rng_);
        if (!y.def.relu)   // head: only the first NN_OUT outputs are used
            for (int co = nn::NN_OUT; co < y.def.cout; ++co)
                std::fill(w.begin() + (size_t)co * 27 * y.def.cin, w.begin() + (size_t)(co + 1) * 27 * y.def.cin, 0.f);
        WSCALL(hipMemcpyAsync(y.w, w.data(), nw * sizeof(float), hipMemcpyHostToDevice, st_));
        nn::weights_to_bf16(st_, y.def.cout, y.def.cin, y.w, y.wb, y.wtb);
    }
    int st = dalloc(&acc_, 4, allocs_);
    if (st) return st;
    WSCALL(hipStreamSynchronize(st_));
    return 0;
}

int WarmStart::param_count() const {
    size_t n = 0;
    for (const Layer &y : L_) n += y.nw() + y.def.cout;
    return (int)n;
}

int WarmStart::get_params(float *host, size_t n) {
    if (n != (size_t)param_count()) return set_error(AMCFD_E_ARG, "parameter buffer has the wrong length");
    size_t o = 0;
    for (Layer &y : L_) {
        WSCALL(hipMemcpyAsync(host + o, y.w, y.nw() * sizeof(float), hipMemcpyDeviceToHost, st_));
        o += y.nw();
        WSCALL(hipMemcpyAsync(host + o, y.b, y.def.cout * sizeof(float), hipMemcpyDeviceToHost, st_));
        o += y.def.cout;
    }
    WSCALL(hipStreamSynchronize(st_));
    return 0;
}

int WarmStart::set_params(const float *host, size_t n) {
    if (n != (size_t)param_count()) return set_error(AMCFD_E_ARG, "parameter buffer has the wrong length");
    size_t o = 0;
    for (Layer &y : L_) {
        WSCALL(hipMemcpyAsync(y.w, host + o, y.nw() * sizeof(float), hipMemcpyHostToDevice, st_));
        o += y.nw();
        WSCALL(hipMemcpyAsync(y.b, host + o, y.def.cout * sizeof(float), hipMemcpyHostToDevice, st_));
        o += y.def.cout;
        nn::weights_to_bf16(st_, y.def.cout, y.def.cin, y.w, y.wb, y.wtb);
        WSCALL(hipMemsetAsync(y.m, 0, y.nw() * sizeof(float), st_));
        WSCALL(hipMemsetAsync(y.v, 0, y.nw() * sizeof(float), st_));
        WSCALL(hipMemsetAsync(y.bm, 0, y.def.cout * sizeof(float), st_));
        WSCALL(hipMemsetAsync(y.bv, 0, y.def.cout * sizeof(float), st_));
    }
    step_t_ = 0;
    WSCALL(hipStreamSynchronize(st_));
    return 0;
}

int WarmStart::save(const char *path) {
    std::vector<float> p(param_count());
    int st = get_params(p.data(), p.size());
    if (st) return st;
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot write ") + path);
    const unsigned int magic = 0x41574d53u, n = (unsigned int)p.size();   // "SMWA"
    const bool ok = std::fwrite(&magic, 4, 1, f) == 1 && std::fwrite(&n, 4, 1, f) == 1 &&
                    std::fwrite(p.data(), sizeof(float), p.size(), f) == p.size();
    std::fclose(f);
    return ok ? 0 : set_error(AMCFD_E_IO, std::string("short write to ") + path);
}

int WarmStart::load(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return set_error(AMCFD_E_IO, std::string("cannot read ") + path);
    unsigned int magic = 0, n = 0;
    std::vector<float> p(param_count());
    const bool ok = std::fread(&magic, 4, 1, f) == 1 && std::fread(&n, 4, 1, f) == 1 && magic == 0x41574d53u &&
                    n == p.size() && std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
    std::fclose(f);
    if (!ok) return set_error(AMCFD_E_PARSE, std::string("not a warm-start weight file of this network: ") + path);
    return set_params(p.data(), p.size());
}

nn::ConvShape WarmStart::shape(int l) const {
    // spatial size at the input of layer l
    static const int in_level[NL] = {0, 0, 1, 2, 3, 2, 1, 0, 0};
    const LayerDef &df = L_[l].def;
    const int lev = in_level[l];
    nn::ConvShape s;
    s.N = N_;
    s.D = D_ >> lev;
    s.H = H_ >> lev;
    s.W = W_ >> lev;
    s.Cin = df.cin;
    s.stride = df.stride;
    s.Do = df.stride == 2 ? s.D / 2 : s.D;
    s.Ho = df.stride == 2 ? s.H / 2 : s.H;
    s.Wo = df.stride == 2 ? s.W / 2 : s.W;
    s.Cout = df.cout;
    return s;
}

int WarmStart::ensure_batch(int N, int D, int H, int W) {
    if (N == N_ && D == D_ && H == H_ && W == W_) return 0;
    if (D % 8 || H % 8 || W % 8 || N < 1) return set_error(AMCFD_E_ARG, "network input extents must be multiples of 8");
    WSCALL(hipStreamSynchronize(st_));
    for (void *p : act_allocs_) WSCALL(hipFree(p));
    act_allocs_.clear();
    N_ = N;
    D_ = D;
    H_ = H;
    W_ = W;
    const size_t P0 = (size_t)N * D * H * W, P1 = P0 / 8, P2 = P0 / 64, P3 = P0 / 512;
    const size_t asz[8] = {P0 * 16, P1 * 32, P2 * 64, P3 * 128, P3 * 128, P2 * 64, P1 * 32, P0 * 16};
    const size_t csz[3] = {P0 * 48, P1 * 96, P2 * 192};   // c0 (full), c1 (/2), c2 (/4)
    int st;
    if ((st = dalloc(&x_, P0 * 16, act_allocs_))) return st;
    for (int a = 0; a < 8; ++a)
        if ((st = dalloc(&a_[a], asz[a], act_allocs_)) || (st = dalloc(&da_[a], asz[a], act_allocs_))) return st;
    for (int c = 0; c < 3; ++c)
        if ((st = dalloc(&c_[c], csz[c], act_allocs_)) || (st = dalloc(&dc_[c], csz[c], act_allocs_))) return st;
    if ((st = dalloc(&dz_, P0 * 48, act_allocs_)) || (st = dalloc(&out_, P0 * 16, act_allocs_)) ||
        (st = dalloc(&dout_, P0 * 16, act_allocs_)) || (st = dalloc(&tgt_, P0 * nn::NN_OUT, act_allocs_)) ||
        (st = dalloc(&mask_, P0, act_allocs_)))
        return st;
    return 0;
}

void WarmStart::forward(hipStream_t st) {
    const int N = N_;
    auto conv = [&](int l, const unsigned short *in, unsigned short *o, float *of) {
        nn::conv_fwd(st, shape(l), in, L_[l].wb, L_[l].b, o, of, L_[l].def.relu);
    };
    conv(0, x_, a_[0], nullptr);
    conv(1, a_[0], a_[1], nullptr);
    conv(2, a_[1], a_[2], nullptr);
    conv(3, a_[2], a_[3], nullptr);
    conv(4, a_[3], a_[4], nullptr);
    nn::up_cat(st, N, D_ / 4, H_ / 4, W_ / 4, 128, 64, a_[4], a_[2], c_[2]);
    conv(5, c_[2], a_[5], nullptr);
    nn::up_cat(st, N, D_ / 2, H_ / 2, W_ / 2, 64, 32, a_[5], a_[1], c_[1]);
    conv(6, c_[1], a_[6], nullptr);
    nn::up_cat(st, N, D_, H_, W_, 32, 16, a_[6], a_[0], c_[0]);
    conv(7, c_[0], a_[7], nullptr);
    conv(8, a_[7], nullptr, out_);
}

// after nn::mse has written dout_: gradients of every layer into gw / gb
int WarmStart::backward(hipStream_t st, float inv_count) {
    (void)inv_count;
    const size_t P0 = (size_t)N_ * D_ * H_ * W_, P1 = P0 / 8, P2 = P0 / 64, P3 = P0 / 512;
    const size_t asz[8] = {P0 * 16, P1 * 32, P2 * 64, P3 * 128, P3 * 128, P2 * 64, P1 * 32, P0 * 16};
    const size_t csz[3] = {P0 * 48, P1 * 96, P2 * 192};
    for (int a = 0; a < 8; ++a) WSCALL(hipMemsetAsync(da_[a], 0, asz[a] * sizeof(float), st));
    for (int c = 0; c < 3; ++c) WSCALL(hipMemsetAsync(dc_[c], 0, csz[c] * sizeof(float), st));
    for (Layer &y : L_) {
        WSCALL(hipMemsetAsync(y.gw, 0, y.nw() * sizeof(float), st));
        WSCALL(hipMemsetAsync(y.gb, 0, y.def.cout * sizeof(float), st));
    }
    auto layer_bwd = [&](int l, const float *dA, const unsigned short *Y, const unsigned short *X, float *dX) {
        const nn::ConvShape s = shape(l);
        const long long n = (long long)s.N * s.Do * s.Ho * s.Wo * s.Cout;
        nn::relu_grad(st, n, s.Cout, dA, Y, L_[l].def.relu, dz_, L_[l].gb);
        nn::conv_bwd(st, s, X, dz_, L_[l].wtb, L_[l].gw, dX);
    };
    const int N = N_;
    layer_bwd(8, dout_, nullptr, a_[7], da_[7]);
    layer_bwd(7, da_[7], a_[7], c_[0], dc_[0]);
    nn::up_cat_bwd(st, N, D_, H_, W_, 32, 16, dc_[0], da_[6], da_[0]);
    layer_bwd(6, da_[6], a_[6], c_[1], dc_[1]);
    nn::up_cat_bwd(st, N, D_ / 2, H_ / 2, W_ / 2, 64, 32, dc_[1], da_[5], da_[1]);
    layer_bwd(5, da_[5], a_[5], c_[2], dc_[2]);
    nn::up_cat_bwd(st, N, D_ / 4, H_ / 4, W_ / 4, 128, 64, dc_[2], da_[4], da_[2]);
    layer_bwd(4, da_[4], a_[4], a_[3], da_[3]);
    layer_bwd(3, da_[3], a_[3], a_[2], da_[2]);
    layer_bwd(2, da_[2], a_[2], a_[1], da_[1]);
    layer_bwd(1, da_[1], a_[1], a_[0], da_[0]);
    layer_bwd(0, da_[0], a_[0], x_, nullptr);
    WSCALL(hipGetLastError());
    return 0;
}

int WarmStart::loss_grad(const float *x, const float *tgt, const float *mask, int N, int D, int H, int W,
                         double *loss, float *grad) {
    std::lock_guard<std::mutex> lock(mu_);
    int st = ensure_batch(N, D, H, W);
    if (st) return st;
    const size_t P0 = (size_t)N * D * H * W;
    float *xf = nullptr;
    WSCALL(hipMalloc(&xf, P0 * 16 * sizeof(float)));
    WSCALL(hipMemcpyAsync(xf, x, P0 * 16 * sizeof(float), hipMemcpyHostToDevice, st_));
    nn::to_bf16(st_, (long long)P0 * 16, xf, x_);
    WSCALL(hipMemcpyAsync(tgt_, tgt, P0 * nn::NN_OUT * sizeof(float), hipMemcpyHostToDevice, st_));
    WSCALL(hipMemcpyAsync(mask_, mask, P0 * sizeof(float), hipMemcpyHostToDevice, st_));
    std::vector<float> mh(mask, mask + P0);
    double cnt = 0.0;
    for (float m : mh) cnt += m;
    const float inv = cnt > 0 ? (float)(1.0 / (cnt * nn::NN_OUT)) : 0.f;
    forward(st_);
    WSCALL(hipMemsetAsync(acc_, 0, sizeof(float), st_));
    nn::mse(st_, (long long)P0, nn::NN_COUT, nn::NN_OUT, out_, tgt_, mask_, inv, dout_, acc_);
    if ((st = backward(st_, inv))) return st;
    float a = 0.f;
    WSCALL(hipMemcpyAsync(&a, acc_, sizeof(float), hipMemcpyDeviceToHost, st_));
    size_t o = 0;
    for (Layer &y : L_) {
        WSCALL(hipMemcpyAsync(grad + o, y.gw, y.nw() * sizeof(float), hipMemcpyDeviceToHost, st_));
        o += y.nw();
        WSCALL(hipMemcpyAsync(grad + o, y.gb, y.def.cout * sizeof(float), hipMemcpyDeviceToHost, st_));
        o += y.def.cout;
    }
    WSCALL(hipStreamSynchronize(st_));
    WSCALL(hipFree(xf));
    *loss = a * inv;
    return 0;
}

int WarmStart::train(int iters, int batch, double lr, double *loss_first, double *loss_last) {
    std::lock_guard<std::mutex> lock(mu_);
    if (nsamp_ < 1) return set_error(AMCFD_E_ARG, "no recorded samples to train on");
    if (iters < 1 || batch < 1) return set_error(AMCFD_E_ARG, "iters and batch must be positive");
    int st = ensure_batch(batch, WD, WH, WW);
    if (st) return st;
    const size_t P = (size_t)WD * WH * WW;
    // the filled slots (samples still pending on some solver are skipped)
    std::vector<int> slots;
    for (int i = 0; i < cap_; ++i)
        if (filled_[i]) slots.push_back(i);
    std::uniform_int_distribution<int> pick(0, (int)slots.size() - 1);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double first = -1.0, last = -1.0;
    for (int it = 0; it < iters; ++it) {
        for (int b = 0; b < batch; ++b) {
            const int s = slots[pick(rng_)];
            WSCALL(hipMemcpyAsync(x_ + b * P * 16, sx_ + (size_t)s * P * 16, P * 16 * sizeof(unsigned short),
                                  hipMemcpyDeviceToDevice, st_));
            WSCALL(hipMemcpyAsync(tgt_ + b * P * nn::NN_OUT, stgt_ + (size_t)s * P * nn::NN_OUT,
                                  P * nn::NN_OUT * sizeof(float), hipMemcpyDeviceToDevice, st_));
            WSCALL(hipMemcpyAsync(mask_ + b * P, smask_ + (size_t)s * P, P * sizeof(float), hipMemcpyDeviceToDevice,
                                  st_));
        }
        const float inv = 1.f / (float)(batch * P * nn::NN_OUT);   // mean over the window (masked cells add 0)
        forward(st_);
        WSCALL(hipMemsetAsync(acc_, 0, sizeof(float), st_));
        nn::mse(st_, (long long)batch * P, nn::NN_COUT, nn::NN_OUT, out_, tgt_, mask_, inv, dout_, acc_);
        if ((st = backward(st_, inv))) return st;
        ++step_t_;
        nn::AdamStep a{(float)lr, b1, b2, eps, (float)(1.0 / (1.0 - std::pow(b1, step_t_))),
                       (float)(1.0 / (1.0 - std::pow(b2, step_t_)))};
        for (Layer &y : L_)
            nn::adam(st_, y.def.cout, y.def.cin, y.w, y.m, y.v, y.gw, y.b, y.bm, y.bv, y.gb, a, y.wb, y.wtb);
        if (it == 0 || it == iters - 1) {
            float l = 0.f;
            WSCALL(hipMemcpyAsync(&l, acc_, sizeof(float), hipMemcpyDeviceToHost, st_));
            WSCALL(hipStreamSynchronize(st_));
            if (it == 0) first = l * inv;
            last = l * inv;
        }
    }
    WSCALL(hipStreamSynchronize(st_));
    if (loss_first) *loss_first = first;
    if (loss_last) *loss_last = last;
    return 0;
}

nn::FeatScale WarmStart::scale(const Solver &s) const {
    const amcfd_params &p = s.params();
    const double peak = p.alasfact * p.alaspow * p.alaseta / (M_PI * p.alasrb * p.alasrb);
    nn::FeatScale f;
    f.t0 = p.tempamb;
    f.tinv = 1e-3;
    f.vinv = 0.1;
    f.qinv = peak > 0 ? 1.0 / peak : 0.0;
    // targets of order one: a molten step changes T near the beam by tens
    // of K (rms; up to ~1e3 K at the spot) and the pressure by ~1e3 Pa (up
    // to ~1e5 Pa), so that neither channel swamps the other in the loss
    f.dtscale = 100.0;
    f.dpscale = 1e4;
    f.dtinv = 1.0 / f.dtscale;
    f.dpinv = 1.0 / f.dpscale;
    return f;
}

// the window: WW x WH cells centred on the beam's nearest node, the top WD
// planes of the active grid
nn::Window WarmStart::window(const Solver &s) const {
    int ni, nj, nk;
    s.dims(&ni, &nj, &nk);
    nn::Window w;
    w.W = WW;
    w.H = WH;
    w.D = WD;
    w.i0 = s.beam_i() - WW / 2;
    w.j0 = s.beam_j() - WH / 2;
    w.k0 = s.active_nk() - WD;
    return w;
}

void WarmStart::attach(Solver *s) {
    std::lock_guard<std::mutex> lock(mu_);
    recs_.emplace(s, Rec{});
}

void WarmStart::detach(Solver *s) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = recs_.find(s);
    if (it == recs_.end()) return;
    (void)hipStreamSynchronize(s->stream());
    (void)hipFree(it->second.told);
    (void)hipFree(it->second.pold);
    recs_.erase(it);
}

int WarmStart::capture_input(Solver &s) {
    const size_t P = (size_t)WD * WH * WW;
    const size_t nn_ = (size_t)s.dev_dims().ni * s.dev_dims().nj * s.dev_dims().nkf;
    Rec r;
    {
        std::lock_guard<std::mutex> lock(mu_);
        if (!sx_) {
            cap_ = 2048;
            int st;
            if ((st = dalloc(&sx_, (size_t)cap_ * P * 16, allocs_)) ||
                (st = dalloc(&stgt_, (size_t)cap_ * P * nn::NN_OUT, allocs_)) ||
                (st = dalloc(&smask_, (size_t)cap_ * P, allocs_)))
                return st;
            WSCALL(hipStreamSynchronize(st_));   // the ring's clears are done before solvers write into it
            filled_.assign(cap_, 0);
        }
        auto it = recs_.find(&s);
        if (it == recs_.end()) return set_error(AMCFD_E_ARG, "solver is not attached to this warm start");
        Rec &m = it->second;
        if (m.n != nn_) {
            // this solver's own buffers: nothing of its is queued on them now
            WSCALL(hipStreamSynchronize(s.stream()));
            (void)hipFree(m.told);
            (void)hipFree(m.pold);
            m.told = m.pold = nullptr;
            m.n = 0;
            WSCALL(hipMalloc(&m.told, nn_ * sizeof(double)));
            WSCALL(hipMalloc(&m.pold, nn_ * sizeof(double)));
            m.n = nn_;
        }
        m.slot = next_;   // a sample begun and never finished leaves its slot unfilled
        next_ = (next_ + 1) % cap_;
        if (filled_[m.slot]) {   // the ring wrapped onto an old sample
            filled_[m.slot] = 0;
            --nsamp_;
        }
        m.win = window(s);
        r = m;
    }
    hipStream_t st = s.stream();
    nn::features(st, s.dev_dims(), r.win, scale(s), s.dev_fields(), sx_ + (size_t)r.slot * P * 16,
                 smask_ + (size_t)r.slot * P);
    WSCALL(hipMemcpyAsync(r.told, s.dev_fields().temp, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st));
    WSCALL(hipMemcpyAsync(r.pold, s.dev_fields().pressure, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st));
    WSCALL(hipGetLastError());
    return 0;
}

int WarmStart::capture_target(Solver &s) {
    const size_t P = (size_t)WD * WH * WW;
    Rec r;
    {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = recs_.find(&s);
        if (it == recs_.end() || it->second.slot < 0) return 0;
        r = it->second;
    }
    nn::target(s.stream(), s.dev_dims(), r.win, scale(s), s.dev_fields().temp, r.told, s.dev_fields().pressure,
               r.pold, stgt_ + (size_t)r.slot * P * nn::NN_OUT);
    WSCALL(hipGetLastError());
    WSCALL(hipStreamSynchronize(s.stream()));
    std::lock_guard<std::mutex> lock(mu_);
    auto it = recs_.find(&s);
    if (it != recs_.end() && it->second.slot == r.slot) {
        it->second.slot = -1;
        filled_[r.slot] = 1;
        nsamp_ = std::min(nsamp_ + 1, cap_);
    }
    return 0;
}

int WarmStart::apply(Solver &s) {
    std::lock_guard<std::mutex> lock(mu_);
    int st = ensure_batch(1, WD, WH, WW);
    if (st) return st;
    hipStream_t sst = s.stream();
    // the network runs on its own stream, ordered after the solver's work
    WSCALL(hipStreamSynchronize(sst));
    const nn::Window w = window(s);
    const nn::FeatScale f = scale(s);
    nn::features(st_, s.dev_dims(), w, f, s.dev_fields(), x_, nullptr);
    forward(st_);
    const amcfd_params &p = s.params();
    nn::GuessConst g{p.tliquid, p.tsolid, p.hlcal, p.hsmelt, p.acpa, p.acpb, p.acpl, tfactor, pfactor};
    nn::apply_guess(st_, s.dev_dims(), w, f, g, out_, s.dev_fields());
    WSCALL(hipGetLastError());
    WSCALL(hipStreamSynchronize(st_));
    return 0;
}

}  // namespace amcfd

// ---------------------------------------------------------------------------
// C ABI (include/amcfd.h)
// ---------------------------------------------------------------------------
struct amcfd_ws {
    amcfd::WarmStart w;
};

extern "C" {

int amcfd_ws_create(int seed, amcfd_ws **out) {
    if (!out) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    *out = nullptr;
    amcfd_ws *h = new (std::nothrow) amcfd_ws;
    if (!h) return amcfd::set_error(AMCFD_E_CAPACITY, "out of host memory");
    int st;
    try {
        st = h->w.create(seed);
    } catch (const std::exception &e) {
        st = amcfd::set_error(AMCFD_E_HIP, e.what());
    }
    if (st) {
        delete h;
        return st;
    }
    *out = h;
    return AMCFD_OK;
}

void amcfd_ws_destroy(amcfd_ws *ws) { delete ws; }

int amcfd_ws_param_count(amcfd_ws *ws) { return ws ? ws->w.param_count() : amcfd::set_error(AMCFD_E_ARG, "null"); }

int amcfd_ws_params(amcfd_ws *ws, float *host, size_t n, int set) {
    if (!ws || !host) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    return set ? ws->w.set_params(host, n) : ws->w.get_params(host, n);
}

int amcfd_ws_loss_grad(amcfd_ws *ws, const float *x, const float *target, const float *mask, int n, int d, int h,
                       int w, double *loss, float *grad) {
    if (!ws || !x || !target || !mask || !loss || !grad) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    return ws->w.loss_grad(x, target, mask, n, d, h, w, loss, grad);
}

int amcfd_ws_train(amcfd_ws *ws, int iters, int batch, double lr, double *loss_first, double *loss_last) {
    if (!ws) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    return ws->w.train(iters, batch, lr, loss_first, loss_last);
}

int amcfd_ws_samples(amcfd_ws *ws) { return ws ? ws->w.samples() : amcfd::set_error(AMCFD_E_ARG, "null"); }

int amcfd_ws_save(amcfd_ws *ws, const char *path) {
    if (!ws || !path) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    return ws->w.save(path);
}

int amcfd_ws_load(amcfd_ws *ws, const char *path) {
    if (!ws || !path) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    return ws->w.load(path);
}

int amcfd_ws_set_factors(amcfd_ws *ws, double tfactor, double pfactor) {
    if (!ws) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    ws->w.tfactor = tfactor;
    ws->w.pfactor = pfactor;
    return AMCFD_OK;
}

}  // extern "C"

namespace amcfd {
int ws_capture_input(WarmStart *w, Solver &s) { return w->capture_input(s); }
int ws_capture_target(WarmStart *w, Solver &s) { return w->capture_target(s); }
int ws_apply(WarmStart *w, Solver &s) { return w->apply(s); }
void ws_attach(WarmStart *w, Solver *s) { w->attach(s); }
void ws_detach(WarmStart *w, Solver *s) { w->detach(s); }
WarmStart *ws_of(amcfd_ws *h) { return h ? &h->w : nullptr; }
}  // namespace amcfd
