// Mixed-precision mode (AMCFD_MODE_MIXED): line sweeps in parallel order,
// FP32 line solves on an FP64 residual.
//
// The reference's SIMPLE iteration smooths each linearised equation with a
// block correction along x and four Gauss-Seidel x-line passes
// (CFDSolverManager.cpp:2270-2362, 2459-2526 on the whole domain,
// 2367-2454 on the melt-pool box).  This mode keeps that structure and the
// reference's coefficients, but
//   - the coefficient arrays of the enthalpy equation are stored in FP32
//     (half the bytes of the dominant kernels);
//   - every pass is a correction in delta form: the residual
//       r = su + sum_nb a_nb phi_nb - ap phi
//     is formed in FP64 from the FP64 state, the line system
//       ap dphi_i - aw dphi_{i-1} - ae dphi_{i+1} = r_i
//     is solved in FP32 and phi += dphi is applied in FP64.  Rounding in the
//     FP32 solve only perturbs the correction, which later passes and
//     iterations remove (iterative refinement): the fixed point is that of
//     the FP64 residual;
//   - the passes run in parallel orders instead of the sequential j/k loop
//     nest: x-lines in red-black (zebra) order over (j, k), one wave per
//     line with the line solved by wave-wide scans (any length up to
//     MX_MAXN); y- and z-lines one lane per line (Jacobi between lines,
//     reading one buffer and writing another), which alternate the implicit
//     direction.
// Block-correction column sums come from per-tile partials that the
// assembly kernel writes; they and the 1-D correction along x are folded in
// FP64 in a fixed order, so a run is reproducible bit for bit.

#include <type_traits>

#include "boxprep.h"
#include "fv.h"

// multiply-adds may fuse here: this mode does not reproduce the reference's
// roundings, and its FP64 residuals are the refinement's accuracy anyway
#pragma clang fp contract(fast)

namespace amcfd {

namespace {

constexpr int MX_MAXN = 4096;   // longest x-line of the wave solver (64 lanes x 64 rows)
constexpr int YZ_MAXL = 300;    // longest y- or z-line of the lane solver (LDS: 64 lanes x 8 B x 300)
constexpr int BF_SPLIT = 16;    // first-level splits of the column-sum fold

__device__ __forceinline__ size_t cix(const Dims &d, int i, int j, int k) {
    return (size_t)k * d.sk + (size_t)j * d.sj + (size_t)i;
}

// exact power-of-two rescale (projective quantities only)
template <typename T>
__device__ __forceinline__ T pow2_inv(T v) {
    const int e = ilogb(fabs(v));
    return (v == T(0) || e > 1000 || e < -1000) ? T(1) : ldexp(T(1), -e);
}

// ------------------------------------------------------------------------
// Wave-wide tridiagonal solve of one line,
//   -w_p x_{p-1} + dg_p x_p - e_p x_{p+1} = r_p,   p = 0..n-1,
// with x outside the line taken as 0 (w_0 and e_{n-1} are not read).  Lane
// l holds rows l*M .. l*M+M-1.  Thomas' forward pivots c_p = e_p/den_p,
// den_p = dg_p - w_p c_{p-1}, and g_p = (r_p + w_p g_{p-1})/den_p follow
// from the projective recurrence
//   (N, D, H)_p = [[0, e, 0], [-w, dg, 0], [0, r, w]]_p (N, D, H)_{p-1}
// with c = N/D, g = H/D and (N, D, H)_{-1} = (0, 1, 0): each lane composes
// its rows' matrices, the lanes combine them with log2(64) shuffle steps,
// and each lane replays its rows from the state entering it (plain Thomas
// steps).  Back substitution x_p = g_p + c_p x_{p+1} is a suffix scan of
// affine maps.  Matrix products are rescaled by powers of two (exact; the
// recurrence is projective), so FP32 neither overflows nor underflows.
// ------------------------------------------------------------------------
template <typename T>
struct Mob {   // [[a, b, 0], [c, d, 0], [f, g, h]]
    T a, b, c, d, f, g, h;
    __device__ __forceinline__ static Mob ident() { return Mob{T(1), T(0), T(0), T(1), T(0), T(0), T(1)}; }
    // this = L * this
    __device__ __forceinline__ void lmul(const Mob &L) {
        const T na = L.a * a + L.b * c, nb = L.a * b + L.b * d;
        const T nc = L.c * a + L.d * c, nd = L.c * b + L.d * d;
        const T nf = L.f * a + L.g * c + L.h * f, ng = L.f * b + L.g * d + L.h * g;
        const T nh = L.h * h;
        a = na; b = nb; c = nc; d = nd; f = nf; g = ng; h = nh;
        const T s = pow2_inv(fabs(d) > fabs(b) ? d : b);
        a *= s; b *= s; c *= s; d *= s; f *= s; g *= s; h *= s;
    }
    // this = R-row step (the row matrix on the left of this)
    __device__ __forceinline__ void row(T w, T dg, T e, T r) {
        const T na = e * c, nb = e * d;
        const T nc = -w * a + dg * c, nd = -w * b + dg * d;
        const T nf = r * c + w * f, ng = r * d + w * g;
        const T nh = w * h;
        a = na; b = nb; c = nc; d = nd; f = nf; g = ng; h = nh;
        const T s = pow2_inv(fabs(d) > fabs(b) ? d : b);
        a *= s; b *= s; c *= s; d *= s; f *= s; g *= s; h *= s;
    }
};

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int off) { return __shfl_up(v, off, 64); }
template <typename T>
__device__ __forceinline__ T shfl_down_t(T v, int off) { return __shfl_down(v, off, 64); }

template <typename T, int M>
__device__ void wave_tridiag(int n, const T (&w)[M], const T (&dg)[M], const T (&e)[M], T (&r)[M], bool guard) {
    const int lane = threadIdx.x & 63;
    Mob<T> S = Mob<T>::ident();
#pragma unroll
    for (int q = 0; q < M; ++q)
        if (lane * M + q < n) S.row(w[q], dg[q], e[q], r[q]);
    // inclusive prefix over lanes: P_l = S_l * ... * S_0.  A step of
    // offset off changes only lanes >= off, so once off reaches the lanes
    // that hold rows the rest change nothing that is read (short lines)
    const int used = (n + M - 1) / M;
    Mob<T> P = S;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        if (off >= used) break;
        Mob<T> L;
        L.a = shfl_up_t(P.a, off); L.b = shfl_up_t(P.b, off); L.c = shfl_up_t(P.c, off);
        L.d = shfl_up_t(P.d, off); L.f = shfl_up_t(P.f, off); L.g = shfl_up_t(P.g, off);
        L.h = shfl_up_t(P.h, off);
        if (lane >= off) {   // P = P * L  (the earlier lanes act first)
            Mob<T> t = L;
            t.lmul(P);
            P = t;
        }
    }
    // state entering this lane: column 2 of the exclusive prefix
    T N = shfl_up_t(P.b, 1), D = shfl_up_t(P.d, 1), H = shfl_up_t(P.g, 1);
    T cp = T(0), gp = T(0);
    if (lane > 0) {
        cp = N / D;
        gp = H / D;
    }
    T c[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
        T den = dg[q] - w[q] * cp;
        if (guard && fabs(den) <= T(1e-12)) den += T(AMCFD_SMALL);
        cp = e[q] / den;
        gp = (r[q] + w[q] * gp) / den;
        c[q] = (lane * M + q < n) ? cp : T(0);
        r[q] = (lane * M + q < n) ? gp : T(0);
    }
    // back substitution: x_p = r_p + c_p x_{p+1}, suffix scan over lanes
    T A = T(1), B = T(0);
#pragma unroll
    for (int q = M - 1; q >= 0; --q) {
        B = c[q] * B + r[q];
        A = c[q] * A;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T Ap = shfl_down_t(A, off), Bp = shfl_down_t(B, off);
        if (lane + off < 64) {
            B = A * Bp + B;
            A = A * Ap;
        }
    }
    const T Bn = shfl_down_t(B, 1);
    T x = (lane == 63) ? T(0) : Bn;
#pragma unroll
    for (int q = M - 1; q >= 0; --q) {
        x = c[q] * x + r[q];
        r[q] = x;
    }
}

// The line in LDS (sw, sd, se: sub-, main and super-diagonal, sr: right-hand
// side, row p at [p]) solved by the wave with M rows per lane; the solution
// replaces sr.  Ends with the wave's writes to sr issued after a barrier.
template <int M>
__device__ __forceinline__ void lds_line_solve(int n, int lane, const float *sw, const float *sd, const float *se,
                                               float *sr, bool guard) {
    float w[M], dg[M], e[M], r[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int p = lane * M + q;
        const bool in = p < n;
        w[q] = in ? sw[p] : 0.f;
        dg[q] = in ? sd[p] : 1.f;
        e[q] = in ? se[p] : 0.f;
        r[q] = in ? sr[p] : 0.f;
    }
    wave_tridiag<float, M>(n, w, dg, e, r, guard);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int p = lane * M + q;
        if (p < n) sr[p] = r[q];
    }
}

// lds_line_solve with the smallest row count per lane (2, 4, ... up to the
// capacity M) that holds the line: the scans' grouping of rows, and so the
// FP32 roundings, depend on the line alone, not on the capacity a launch
// was built for (grouped launches serve lines of several sizes)
template <int M>
__device__ __forceinline__ void lds_line_solve_fit(int n, int lane, const float *sw, const float *sd,
                                                   const float *se, float *sr, bool guard) {
    if constexpr (M > 2) {
        if (n <= 64 * (M / 2)) {
            lds_line_solve_fit<M / 2>(n, lane, sw, sd, se, sr, guard);
            return;
        }
    }
    lds_line_solve<M>(n, lane, sw, sd, se, sr, guard);
}

template <typename TC>
struct MxCoefs {
    const TC *ap, *an, *as, *ae, *aw, *at, *ab;
    const double *su;
};

// ------------------------------------------------------------------------
// x-lines of one colour: (j + k) % 2 == colour over [jbeg, jend] x [kbeg,
// kend], rows i in [ibeg, iend]; ends ibeg-1 and iend+1 are fixed.  One
// wave per line; global loads are row-contiguous (lane-cyclic), the solve
// is lane-blocked, LDS transposes between the two.  Lines of one colour
// read only lines of the other, so the pass is race-free and in place.
// ------------------------------------------------------------------------
template <typename TC, int M>
struct XlineK {
    static constexpr int family = (std::is_same<TC, float>::value ? 11 : 12), rank = M;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, MxCoefs<TC> cf, double *phi, int ibeg, int iend, int jbeg,
                                                  int jend, int kbeg, int kend, int colour, int guard,
                                                  const int *__restrict__ skip, const int *__restrict__ dbox) {
    __shared__ float sw[64 * M], sd[64 * M], se[64 * M], sr[64 * M];
    if (skip && *skip) return;
    if (dbox) {   // the melt-pool box the melt-pool kernel left on the device
        ibeg = dbox[0] + 1;
        iend = dbox[1] - 1;
        jbeg = dbox[2];
        jend = dbox[3];
        kbeg = dbox[4];
        if (iend - ibeg + 1 > 64 * M || iend < ibeg || jend < jbeg || kend < kbeg) return;
    }
    const int lane = threadIdx.x;
    const int n = iend - ibeg + 1, nj = jend - jbeg + 1, nk = kend - kbeg + 1;
    const int cm = (nj + 1) / 2;
    const long long nlines = (long long)nk * cm;
    // workgroups go to the 8 XCDs round-robin by id; give each XCD one
    // contiguous run of lines, so that the lines sharing a neighbour line
    // (j +- 2 and k +- 2 of the same colour) read it through one L2
    const int G = vb.gx, q8 = G / 8, r8 = G % 8, x8 = vb.x % 8, i8 = vb.x / 8;
    const int t = !vb.hw ? vb.x : (x8 < r8) ? x8 * (q8 + 1) + i8 : r8 * (q8 + 1) + (x8 - r8) * q8 + i8;
    const long long chunk = (nlines + G - 1) / G;
    const long long u0 = (long long)t * chunk, u1 = min(u0 + chunk, nlines);
    for (long long u = u0; u < u1; ++u) {
        const int kk = (int)(u / cm), jj = (int)(u % cm);
        const int k = kbeg + kk;
        const int j = jbeg + ((colour - k - jbeg) & 1) + 2 * jj;
        if (j > jend) continue;   // uniform over the wave
        const size_t row = (size_t)k * d.sk + (size_t)j * d.sj + ibeg;
        double hp[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int ii = lane + 64 * q;
            hp[q] = 0.0;
            if (ii < n) {
                const size_t c = row + ii;
                const double h = phi[c];
                const double ap = (double)cf.ap[c], aw = (double)cf.aw[c], ae = (double)cf.ae[c];
                const double r = cf.su[c] + aw * phi[c - 1] + ae * phi[c + 1] + (double)cf.an[c] * phi[c + d.sj] +
                                 (double)cf.as[c] * phi[c - d.sj] + (double)cf.at[c] * phi[c + d.sk] +
                                 (double)cf.ab[c] * phi[c - d.sk] - ap * h;
                hp[q] = h;
                sw[ii] = (float)aw;
                sd[ii] = (float)ap;
                se[ii] = (float)ae;
                sr[ii] = (float)r;
            }
        }
        __syncthreads();
        // the fewest rows per lane that hold the line (box lines are much
        // shorter than the capacity the domain sets)
        lds_line_solve_fit<M>(n, lane, sw, sd, se, sr, guard != 0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int ii = lane + 64 * q;
            if (ii < n) phi[row + ii] = hp[q] + (double)sr[ii];
        }
        __syncthreads();
    }
}
};
template <typename TC, int M>
__global__ __launch_bounds__(64) void k_xline_mx(Dims d, MxCoefs<TC> cf, double *phi, int ibeg, int iend, int jbeg,
                                                  int jend, int kbeg, int kend, int colour, int guard,
                                                  const int *__restrict__ skip, const int *__restrict__ dbox) {
    XlineK<TC, M>::run(vb_hw(), d, cf, phi, ibeg, iend, jbeg, jend, kbeg, kend, colour, guard, skip, dbox);
}

// ------------------------------------------------------------------------
// Melt-pool box x-lines, all red-black passes of one variable in one
// persistent launch.  Pass p writes its lines into pass buffer p (every box
// cell of every buffer holds BX_PENDING beforehand, BoxPrepK); the last
// pass also writes the field.  A red line of pass p reads its own and its
// black neighbours' values from buffer p-1 (the field for p = 0), a black
// line its own from buffer p-1 and its red neighbours from buffer p; cells
// outside the lines (the box's faces, the lines' ends) come from the field,
// which no task writes before every reader of those cells is done.  Each
// value is polled until it is no longer BX_PENDING, with device-coherent
// relaxed atomic loads, so a value is its own ready flag.  Tasks (pass,
// colour, line) are claimed from a counter in that order, and a task reads
// only values of tasks claimed before it by running workgroups, so the
// schedule cannot deadlock however many workgroups are resident.  Every
// line sees the same values as in the launch-per-half-pass form
// (k_xline_mx), so the results are the same bits.
// ------------------------------------------------------------------------
constexpr long long BX_SPIN_LIMIT = 1LL << 20;                      // bounded wait (~ a second)

__device__ __forceinline__ double bx_ld(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void bx_st(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool bx_pending(double v) {
    return __builtin_bit_cast(unsigned long long, v) == BX_PENDING;
}
// v, already read from p, re-read while pending.  False once the bounded
// wait expired or another wait of the solver has given up.
__device__ __forceinline__ bool bx_await(const double *p, double &v, const int *error) {
    long long spins = 0;
    while (bx_pending(v)) {
        v = bx_ld(p);
        if (++spins > BX_SPIN_LIMIT) return false;
        if ((spins & 255) == 0 && __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return false;
    }
    return true;
}

// the cells of one box line task and where its values come from
struct BoxRow {
    const double *own, *phi, *src_n, *src_s, *src_t, *src_b;
    size_t row;
    int n, lane;
};
// the registers of v are live here: their loads are issued before this point
// and not moved after it (an empty statement that reads them, no code)
template <typename T>
__device__ __forceinline__ void live_one(T v) {
    asm volatile("" ::"v"(v));
}
template <typename... T>
__device__ __forceinline__ void live_here(T... v) {
    (live_one(v), ...);
}
// K rows per lane of a box line: every row's coefficient and value loads
// issued together, the tridiagonal coefficients to LDS, then the values
// still pending polled and the residual formed.  The coefficients hold no
// pass's values, so their loads overlap the wait for the neighbours'; the
// results are those of loading row by row.
template <int K>
__device__ __forceinline__ bool box_rows(const Dims &d, const MxCoefs<double> &cf, const BoxRow &br, double *hp,
                                         float *sw, float *sd, float *se, float *sr, int *error) {
    double ap[K], aw[K], ae[K], an[K], as[K], at[K], ab[K], su[K];
    double h0[K], hw[K], he[K], hn[K], hs[K], ht[K], hb[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int ii = br.lane + 64 * q;
        if (ii < br.n) {
            const size_t c = br.row + ii;
            ap[q] = cf.ap[c]; aw[q] = cf.aw[c]; ae[q] = cf.ae[c]; an[q] = cf.an[c];
            as[q] = cf.as[c]; at[q] = cf.at[c]; ab[q] = cf.ab[c]; su[q] = cf.su[c];
            h0[q] = bx_ld(br.own + c);
            hw[q] = bx_ld((ii > 0 ? br.own : br.phi) + c - 1);
            he[q] = bx_ld((ii < br.n - 1 ? br.own : br.phi) + c + 1);
            hn[q] = bx_ld(br.src_n + c + d.sj);
            hs[q] = bx_ld(br.src_s + c - d.sj);
            ht[q] = bx_ld(br.src_t + c + d.sk);
            hb[q] = bx_ld(br.src_b + c - d.sk);
        }
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int ii = br.lane + 64 * q;
        if (ii < br.n) {
            sw[ii] = (float)aw[q];
            sd[ii] = (float)ap[q];
            se[ii] = (float)ae[q];
            live_here(an[q], as[q], at[q], ab[q], su[q]);
        }
    }
    // every value still pending is reloaded at once, round after round, so
    // the residual can be formed one round trip after the last neighbour's
    // store lands, in whatever order the neighbours land (bounded as
    // bx_await is)
    bool ok = true;
    for (long long spins = 0;;) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const int ii = br.lane + 64 * q;
            if (ii < br.n) {
                const size_t c = br.row + ii;
                if (bx_pending(h0[q])) { any = true; h0[q] = bx_ld(br.own + c); }
                if (bx_pending(hw[q])) { any = true; hw[q] = bx_ld((ii > 0 ? br.own : br.phi) + c - 1); }
                if (bx_pending(he[q])) { any = true; he[q] = bx_ld((ii < br.n - 1 ? br.own : br.phi) + c + 1); }
                if (bx_pending(hn[q])) { any = true; hn[q] = bx_ld(br.src_n + c + d.sj); }
                if (bx_pending(hs[q])) { any = true; hs[q] = bx_ld(br.src_s + c - d.sj); }
                if (bx_pending(ht[q])) { any = true; ht[q] = bx_ld(br.src_t + c + d.sk); }
                if (bx_pending(hb[q])) { any = true; hb[q] = bx_ld(br.src_b + c - d.sk); }
            }
        }
        if (!any) break;
        if (++spins > BX_SPIN_LIMIT ||
            ((spins & 255) == 0 && __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
            ok = false;
            break;
        }
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int ii = br.lane + 64 * q;
        if (ii < br.n) {
            const double r = su[q] + aw[q] * hw[q] + ae[q] * he[q] + an[q] * hn[q] + as[q] * hs[q] +
                             at[q] * ht[q] + ab[q] * hb[q] - ap[q] * h0[q];
            hp[q] = h0[q];
            sr[ii] = (float)r;
        }
    }
    return ok;
}

__global__ __launch_bounds__(256) void k_box_prep(Dims d, double *bufs, size_t bstride, int passes, int kend,
                                                  unsigned int *claim, const int *__restrict__ skip,
                                                  const int *__restrict__ dbox) {
    BoxPrepK::run(vb_hw(), d, bufs, bstride, passes, kend, claim, skip, dbox);
}

template <int M>
struct BoxLinesK {
    static constexpr int family = 14, rank = M;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, MxCoefs<double> cf, double *phi, double *bufs,
                                                   size_t bstride, int passes, int kend, int guard,
                                                   unsigned int *claim, int *error, const int *__restrict__ skip,
                                                   const int *__restrict__ dbox) {
    __shared__ float sw[64 * M], sd[64 * M], se[64 * M], sr[64 * M];
    __shared__ int s_task;
    if (skip && *skip) return;
    BoxRange b;
    if (!box_range(dbox, kend, 64 * M, b)) return;
    const int lane = threadIdx.x;
    const int n = b.iend - b.ibeg + 1, nj = b.jend - b.jbeg + 1, nk = b.kend - b.kbeg + 1;
    const int cm = (nj + 1) / 2;
    const int per_half = nk * cm, total = 2 * passes * per_half;
    bool ok = true;
    for (;;) {
        if (lane == 0) {
            const unsigned int t = atomicAdd(claim, 1u);
            s_task = t < (unsigned int)total ? (int)t : -1;
        }
        __syncthreads();
        const int t = s_task;
        __syncthreads();
        if (t < 0) break;
        const int h = t / per_half, u = t % per_half;
        const int pass = h >> 1, colour = h & 1;
        const int k = b.kbeg + u / cm;
        const int j = b.jbeg + ((colour - k - b.jbeg) & 1) + 2 * (u % cm);
        if (j > b.jend) continue;   // uniform over the wave
        const double *own = pass == 0 ? phi : bufs + (size_t)(pass - 1) * bstride;
        const double *nbr = colour == 0 ? own : bufs + (size_t)pass * bstride;
        double *dst = bufs + (size_t)pass * bstride;
        const bool last = pass == passes - 1;
        // the line's neighbours: solved lines of the other colour inside the box, fixed values outside
        const double *src_n = j < b.jend ? nbr : phi, *src_s = j > b.jbeg ? nbr : phi;
        const double *src_t = k < b.kend ? nbr : phi, *src_b = k > b.kbeg ? nbr : phi;
        const size_t row = (size_t)k * d.sk + (size_t)j * d.sj + b.ibeg;
        double hp[M];
#pragma unroll
        for (int q = 0; q < M; ++q) hp[q] = 0.0;
        const BoxRow br{own, phi, src_n, src_s, src_t, src_b, row, n, lane};
        // lines of up to 256 cells (the box's, in practice): every load of
        // every row at once; longer ones row by row
        if constexpr (M >= 2) {
            if (n <= 128) {
                ok &= box_rows<2>(d, cf, br, hp, sw, sd, se, sr, error);
                goto loaded;
            }
        }
        if constexpr (M >= 4) {
            if (n <= 256) {
                ok &= box_rows<4>(d, cf, br, hp, sw, sd, se, sr, error);
                goto loaded;
            }
        }
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int ii = lane + 64 * q;
            if (ii < n) {
                const size_t c = row + ii;
                const double *pw = (ii > 0 ? own : phi) + c - 1, *pe = (ii < n - 1 ? own : phi) + c + 1;
                // every load first, then the values still pending are polled
                double h0 = bx_ld(own + c), hw = bx_ld(pw), he = bx_ld(pe);
                double hn = bx_ld(src_n + c + d.sj), hs = bx_ld(src_s + c - d.sj);
                double ht = bx_ld(src_t + c + d.sk), hb = bx_ld(src_b + c - d.sk);
                ok &= bx_await(own + c, h0, error);
                ok &= bx_await(pw, hw, error);
                ok &= bx_await(pe, he, error);
                ok &= bx_await(src_n + c + d.sj, hn, error);
                ok &= bx_await(src_s + c - d.sj, hs, error);
                ok &= bx_await(src_t + c + d.sk, ht, error);
                ok &= bx_await(src_b + c - d.sk, hb, error);
                const double ap = cf.ap[c], aw = cf.aw[c], ae = cf.ae[c];
                const double r = cf.su[c] + aw * hw + ae * he + cf.an[c] * hn + cf.as[c] * hs + cf.at[c] * ht +
                                 cf.ab[c] * hb - ap * h0;
                hp[q] = h0;
                sw[ii] = (float)aw;
                sd[ii] = (float)ap;
                se[ii] = (float)ae;
                sr[ii] = (float)r;
            }
        }
    loaded:
        // one lane's expired wait ends the whole wave (the task loop stays uniform)
        ok = !__any(!ok);
        if (!ok) break;
        __syncthreads();
        // the fewest rows per lane that hold the line (the melt-pool box's
        // lines are much shorter than the domain's)
        lds_line_solve_fit<M>(n, lane, sw, sd, se, sr, guard != 0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int ii = lane + 64 * q;
            if (ii < n) {
                const double v = hp[q] + (double)sr[ii];
                bx_st(dst + row + ii, v);
                if (last) phi[row + ii] = v;
            }
        }
        __syncthreads();
    }
    if (!ok && lane == 0) atomicOr(error, 1);
}
};
template <int M>
__global__ __launch_bounds__(64) void k_box_lines(Dims d, MxCoefs<double> cf, double *phi, double *bufs,
                                                   size_t bstride, int passes, int kend, int guard,
                                                   unsigned int *claim, int *error, const int *__restrict__ skip,
                                                   const int *__restrict__ dbox) {
    BoxLinesK<M>::run(vb_hw(), d, cf, phi, bufs, bstride, passes, kend, guard, claim, error, skip, dbox);
}

// ------------------------------------------------------------------------
// y-lines (DIR 1, along j) or z-lines (DIR 2, along k), one lane per line,
// lanes along i.  Reads hin everywhere and writes hout on the lines' cells
// (Jacobi between lines: no line reads another's update).  The forward
// Thomas values of every row go to LDS, then back substitution writes
// hout = hin + dphi.  blockIdx.y is the coordinate the lines do not run
// along (k for y-lines, j for z-lines).  Few lines per device means few
// waves, so the next row's loads are in flight while the lane works on the
// current one, and the line's own neighbours come from the rows held.
// ------------------------------------------------------------------------
template <typename TC>
struct YzRow {
    float ap, an, as, ae, aw, at, ab;
    double su, h, hw, he, hx1, hx2;   // own, i-1, i+1, the two cross-line neighbours
};

template <typename TC, int DIR>
__device__ __forceinline__ void yz_load(const Dims &d, const MxCoefs<TC> &cf, const double *__restrict__ hin,
                                        size_t c, YzRow<TC> &r) {
    r.ap = (float)cf.ap[c]; r.an = (float)cf.an[c]; r.as = (float)cf.as[c]; r.ae = (float)cf.ae[c];
    r.aw = (float)cf.aw[c]; r.at = (float)cf.at[c]; r.ab = (float)cf.ab[c];
    r.su = cf.su[c];
    r.h = hin[c];
    r.hw = hin[c - 1];
    r.he = hin[c + 1];
    const size_t x = DIR == 1 ? d.sk : d.sj;   // the direction across the line
    r.hx1 = hin[c + x];
    r.hx2 = hin[c - x];
}

template <typename TC, int DIR>
__global__ __launch_bounds__(64) void k_yzline_mx(Dims d, MxCoefs<TC> cf, const double *__restrict__ hin,
                                                   double *__restrict__ hout, int ibeg, int iend, int jbeg, int jend,
                                                   int kbeg, int kend, const int *__restrict__ skip) {
    extern __shared__ float sbuf[];   // [L][64] c, then [L][64] g
    if (skip && *skip) return;
    const int lane = threadIdx.x;
    const int i = min(ibeg + (int)blockIdx.x * 64 + lane, iend);   // idle lanes repeat the last line, unstored
    const bool act = ibeg + (int)blockIdx.x * 64 + lane <= iend;
    const int L = DIR == 1 ? jend - jbeg + 1 : kend - kbeg + 1;
    const size_t stride = DIR == 1 ? d.sj : d.sk;
    const size_t base = DIR == 1 ? cix(d, i, jbeg, kbeg + blockIdx.y) : cix(d, i, jbeg + blockIdx.y, kbeg);
    float *sc = sbuf, *sg = sbuf + (size_t)L * 64;
    float cp = 0.f, gp = 0.f;
    YzRow<TC> cur, nxt;
    yz_load<TC, DIR>(d, cf, hin, base, cur);
    double hprev = hin[base - stride];   // the fixed end below the line
    for (int s = 0; s < L; ++s) {
        const bool more = s + 1 < L;
        if (more) yz_load<TC, DIR>(d, cf, hin, base + (size_t)(s + 1) * stride, nxt);
        // the next row's own value, or the fixed end above the line
        const double hnext = more ? nxt.h : hin[base + (size_t)L * stride];
        double hN, hS, hT, hB;
        if (DIR == 1) { hN = hnext; hS = hprev; hT = cur.hx1; hB = cur.hx2; }
        else { hT = hnext; hB = hprev; hN = cur.hx1; hS = cur.hx2; }
        const double res = cur.su + (double)cur.aw * cur.hw + (double)cur.ae * cur.he + (double)cur.an * hN +
                           (double)cur.as * hS + (double)cur.at * hT + (double)cur.ab * hB - (double)cur.ap * cur.h;
        const float sub = DIR == 1 ? cur.as : cur.ab;
        const float sup = DIR == 1 ? cur.an : cur.at;
        const float den = cur.ap - sub * cp;
        cp = sup / den;
        gp = ((float)res + sub * gp) / den;
        sc[s * 64 + lane] = cp;
        sg[s * 64 + lane] = gp;
        hprev = cur.h;
        if (more) cur = nxt;
    }
    float x = 0.f;
    for (int s = L - 1; s >= 0; --s) {
        x = sg[s * 64 + lane] + sc[s * 64 + lane] * x;
        const size_t c = base + (size_t)s * stride;
        if (act) hout[c] = hin[c] + (double)x;
    }
}

// ------------------------------------------------------------------------
// block correction along x from the assembly's per-tile column partials
// part[q][t][i] (q = bl, blp, blm, blc; t = (j, k) tile): a fixed-order fold
// in two levels, then the 1-D system
//   bl_i d_i - blm_i d_{i-1} - blp_i d_{i+1} = blc_i
// (CFDSolverManager.cpp:2334-2349) solved by one wave, then phi += d_i.
// ------------------------------------------------------------------------
struct BlkFoldK {
    static constexpr int family = 8, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, int ni, int ntile, const double *__restrict__ part, double *out,
                                                  const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int i = vb.x * 64 + threadIdx.x, q = vb.y, s = vb.z;
    if (i >= ni) return;
    const double *p = part + (size_t)q * ntile * ni + i;
    double a = 0.0;
    // the tiles in order, eight tiles' loads in flight at a time
    constexpr int U = 8;
    int t = s;
    for (; t + (U - 1) * BF_SPLIT < ntile; t += U * BF_SPLIT) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = p[(size_t)(t + u * BF_SPLIT) * ni];
#pragma unroll
        for (int u = 0; u < U; ++u) a += v[u];
    }
    for (; t < ntile; t += BF_SPLIT) a += p[(size_t)t * ni];
    out[((size_t)q * BF_SPLIT + s) * ni + i] = a;
}
};
__global__ __launch_bounds__(64) void k_blk_fold(int ni, int ntile, const double *__restrict__ part, double *out,
                                                  const int *__restrict__ skip) {
    BlkFoldK::run(vb_hw(), ni, ntile, part, out, skip);
}

constexpr int BS_NT = 1024;
constexpr int BS_MAXNI = 4100;   // ni of the longest x-line, rounded up

// the 1-D system along x with the fewest rows per lane that hold it (the
// roundings depend on the system alone, whatever capacity M a launch has)
template <int M>
__device__ __forceinline__ void blk_solve_rows(int n, int lane, const double (*sq)[BS_MAXNI], double *delphi) {
    double w[M], dg[M], e[M], r[M];
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int p = lane * M + q, i = p + 1;
        const bool in = p < n;
        dg[q] = in ? sq[0][i] : 1.0;
        e[q] = in ? sq[1][i] : 0.0;
        w[q] = in ? sq[2][i] : 0.0;
        r[q] = in ? sq[3][i] : 0.0;
    }
    wave_tridiag<double, M>(n, w, dg, e, r, false);
#pragma unroll
    for (int q = 0; q < M; ++q) {
        const int p = lane * M + q;
        if (p < n) delphi[p + 1] = r[q];
    }
}
template <int M>
__device__ __forceinline__ void blk_solve_fit(int n, int lane, const double (*sq)[BS_MAXNI], double *delphi) {
    if constexpr (M > 2) {
        if (n <= 64 * (M / 2)) {
            blk_solve_fit<M / 2>(n, lane, sq, delphi);
            return;
        }
    }
    blk_solve_rows<M>(n, lane, sq, delphi);
}

template <int M>
struct BlkSolveK {
    static constexpr int family = 9, rank = M;
    static __device__ __forceinline__ void run(const VB &vb, int nim2, int ni, const double *__restrict__ f1,
                                                        double *delphi, const int *__restrict__ skip) {
    __shared__ double sq[4][BS_MAXNI];
    if (skip && *skip) return;
    // second fold level, all threads: sq[q][i] = sum over the splits
    for (int e = threadIdx.x; e < 4 * ni; e += BS_NT) {
        const int q = e / ni, i = e % ni;
        double a = 0.0;
#pragma unroll
        for (int sp = 0; sp < BF_SPLIT; ++sp) a += f1[((size_t)q * BF_SPLIT + sp) * ni + i];
        sq[q][i] = a;
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    blk_solve_fit<M>(nim2, threadIdx.x, sq, delphi);
}
};
template <int M>
__global__ __launch_bounds__(BS_NT) void k_blk_solve_mx(int nim2, int ni, const double *__restrict__ f1,
                                                        double *delphi, const int *__restrict__ skip) {
    BlkSolveK<M>::run(vb_hw(), nim2, ni, f1, delphi, skip);
}

struct AddDelphiMxK {
    static constexpr int family = 10, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, const double *__restrict__ delphi, double *phi,
                                                      const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int i = 1 + vb.x * 64 + threadIdx.x;
    const int j = 1 + vb.y * 4 + threadIdx.y;
    if (i > d.nim2 || j > d.njm2) return;
    const double dp = delphi[i];
    // the four planes' loads before any store, so they are in flight together
    const int k0 = 1 + vb.z * 4, nk = min(4, d.nkm2 - k0 + 1);
    double v[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        if (kk < nk) v[kk] = phi[cix(d, i, j, k0 + kk)];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        if (kk < nk) phi[cix(d, i, j, k0 + kk)] = v[kk] + dp;
}
};
__global__ __launch_bounds__(256) void k_add_delphi_mx(Dims d, const double *__restrict__ delphi, double *phi,
                                                      const int *__restrict__ skip) {
    AddDelphiMxK::run(vb_hw(), d, delphi, phi, skip);
}

template <typename TC>
void launch_xline_t(hipStream_t st, const Dims &d, const MxCoefs<TC> &cf, double *phi, int ibeg, int iend, int jbeg,
                    int jend, int kbeg, int kend, int colour, int guard, const int *skip, const int *dbox = nullptr,
                    int nblocks = 0, int line_batch = 1) {
    // with dbox the line range is read on the device: the host range is an
    // upper bound (the whole interior) and fixes the kernel's row capacity
    const int n = iend - ibeg + 1;
    const long long nlines = (long long)(kend - kbeg + 1) * ((jend - jbeg + 2) / 2);
    const int nb = dbox ? nblocks : (int)std::min<long long>((nlines + line_batch - 1) / line_batch, 16384);
#define XL(MM) dispatch<XlineK<TC, MM>, 64>(k_xline_mx<TC, MM>, dim3(nb), dim3(64), 0, st, d, cf, phi, ibeg, iend, jbeg, jend, \
                                  kbeg, kend, colour, guard, skip, dbox)
    if (n <= 128) XL(2);
    else if (n <= 256) XL(4);
    else if (n <= 512) XL(8);
    else if (n <= 1024) XL(16);
    else if (n <= 2048) XL(32);
    else XL(64);
#undef XL
}

template <typename TC>
void launch_yzline_t(hipStream_t st, const Dims &d, const MxCoefs<TC> &cf, const double *hin, double *hout, int dir,
                     int ibeg, int iend, int jbeg, int jend, int kbeg, int kend, const int *skip) {
    const int L = dir == 1 ? jend - jbeg + 1 : kend - kbeg + 1;
    const int other = dir == 1 ? kend - kbeg + 1 : jend - jbeg + 1;
    const dim3 grid((iend - ibeg + 64) / 64, other);
    const size_t lds = (size_t)L * 64 * 2 * sizeof(float);
    // more dynamic LDS than the default per-launch limit is allowed once per
    // kernel (the line length was checked against the device's limit)
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)k_yzline_mx<TC, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)((size_t)YZ_MAXL * 64 * 2 * sizeof(float)));
        (void)hipFuncSetAttribute((const void *)k_yzline_mx<TC, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)((size_t)YZ_MAXL * 64 * 2 * sizeof(float)));
        attr_set = true;
    }
    if (LaunchRecorder *r = tls_recorder()) {   // y/z passes have no grouped form
        r->unsupported = true;
        return;
    }
    if (dir == 1)
        hipLaunchKernelGGL((k_yzline_mx<TC, 1>), grid, dim3(64), lds, st, d, cf, hin, hout, ibeg, iend, jbeg, jend,
                           kbeg, kend, skip);
    else
        hipLaunchKernelGGL((k_yzline_mx<TC, 2>), grid, dim3(64), lds, st, d, cf, hin, hout, ibeg, iend, jbeg, jend,
                           kbeg, kend, skip);
}

}  // namespace

int mixed_max_line() { return MX_MAXN; }
int mixed_max_yz() { return YZ_MAXL; }
size_t mixed_yz_lds(int len) { return (size_t)len * 64 * 2 * sizeof(float); }
int mixed_blk_split() { return BF_SPLIT; }

void launch_mixed_block_correction(hipStream_t st, const Dims &d, const MxFields &m, double *phi, const int *skip) {
    dispatch<BlkFoldK, 64>(k_blk_fold, dim3((d.ni + 63) / 64, 4, BF_SPLIT), dim3(64), 0, st, d.ni, mixed_tile_count(d),
                       m.blkpart, m.blkfold, skip);
    const int n = d.nim2;
#define BS(MM) dispatch<BlkSolveK<MM>, BS_NT>(k_blk_solve_mx<MM>, dim3(1), dim3(BS_NT), 0, st, d.nim2, d.ni, m.blkfold, m.delphi, skip)
    if (n <= 128) BS(2);
    else if (n <= 256) BS(4);
    else if (n <= 512) BS(8);
    else if (n <= 1024) BS(16);
    else if (n <= 2048) BS(32);
    else BS(64);
#undef BS
    dispatch<AddDelphiMxK, 256>(k_add_delphi_mx, dim3((d.nim2 + 63) / 64, (d.njm2 + 3) / 4, (d.nkm2 + 3) / 4), dim3(64, 4), 0,
                       st, d, m.delphi, phi, skip);
}

void launch_mixed_xline(hipStream_t st, const Dims &d, const MxFields &m, double *phi, int ibeg, int iend, int jbeg,
                        int jend, int kbeg, int kend, int colour, const int *skip, int line_batch) {
    MxCoefs<float> cf{m.ap, m.an, m.as, m.ae, m.aw, m.at, m.ab, m.su};
    launch_xline_t<float>(st, d, cf, phi, ibeg, iend, jbeg, jend, kbeg, kend, colour, 0, skip, nullptr, 0,
                          std::max(1, line_batch));
}

void launch_mixed_yzline(hipStream_t st, const Dims &d, const MxFields &m, const double *hin, double *hout, int dir,
                         int ibeg, int iend, int jbeg, int jend, int kbeg, int kend, const int *skip) {
    MxCoefs<float> cf{m.ap, m.an, m.as, m.ae, m.aw, m.at, m.ab, m.su};
    launch_yzline_t<float>(st, d, cf, hin, hout, dir, ibeg, iend, jbeg, jend, kbeg, kend, skip);
}

// all red-black passes of one box variable in one persistent launch
int launch_box_passes(hipStream_t st, const Dims &d, DevFields f, double *phi, double *bufs, size_t bstride,
                      int passes, unsigned int *claim, int *error, const int *skip, const int *dbox, int nblocks,
                      bool prep) {
    MxCoefs<double> cf{f.ap, f.an, f.as, f.ae, f.aw, f.at, f.ab, f.su};
    if (prep)
        dispatch<BoxPrepK, 256>(k_box_prep, dim3(nblocks), dim3(256), 0, st, d, bufs, bstride, passes, d.nkm2, claim,
                                skip, dbox);
    const int n = d.nim2;   // an upper bound of the box's line length
#define BL(MM) dispatch<BoxLinesK<MM>, 64>(k_box_lines<MM>, dim3(nblocks), dim3(64), 0, st, d, cf, phi, bufs, bstride, \
                                  passes, d.nkm2, 1, claim, error, skip, dbox)
    if (n <= 128) BL(2);
    else if (n <= 256) BL(4);
    else if (n <= 512) BL(8);
    else if (n <= 1024) BL(16);
    else if (n <= 2048) BL(32);
    else BL(64);
#undef BL
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// melt-pool box lines (FP64 coefficients of the momentum / pressure
// equations, FP32 solve on the FP64 residual, guard as the reference's)
void launch_box_xline(hipStream_t st, const Dims &d, DevFields f, double *phi, int ibeg, int iend, int jbeg, int jend,
                      int kbeg, int kend, int colour, const int *skip, const int *dbox, int nblocks) {
    MxCoefs<double> cf{f.ap, f.an, f.as, f.ae, f.aw, f.at, f.ab, f.su};
    launch_xline_t<double>(st, d, cf, phi, ibeg, iend, jbeg, jend, kbeg, kend, colour, 1, skip, dbox, nblocks);
}

}  // namespace amcfd
