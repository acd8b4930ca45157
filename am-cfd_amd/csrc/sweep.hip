// Gauss-Seidel-ordered x-line sweeps on the GPU.
//
// The reference sweeps  for ks in {0,1}: for k = kend..kbeg: for js in {0,1}:
// for j = jbeg..jend: Thomas along i  (CFDSolverManager.cpp:2415-2453 on the
// melt-pool box, 2492-2525 on the whole domain).  Each line's right-hand side
// reads the lines (k+-1, j) and (k, j+-1) as they stand at that point of the
// sequence, so the sweep is Gauss-Seidel in j and k.  Run single-threaded
// that order is the reference's definition of the result.
//
// Here every line solve is a task; task (pass p = 2ks+js, k, j) needs
//   line (k, j-1) as left by pass p,
//   line (k, j+1) as left by pass p-1               (the sweep's input if p = 0),
//   line (k+1, j) as left by pass 2ks+1,
//   line (k-1, j) as left by pass 1                 (the input if ks = 0),
// which is exactly what the sequential loop nest reads.  Each pass writes its
// lines into a buffer of its own (pass[p]), so every value a task reads is
// written once per sweep and nothing is overwritten while it may still be
// read; pass 3 also writes the field in place, which is safe because every
// task that reads a line's input value precedes that line's pass-3 task in
// the dependency order.  The pass buffers are filled with a signalling-NaN
// pattern that no arithmetic produces before the sweep, and a task polls its
// neighbour values until the pattern is gone: the data is its own flag.
// Tasks are claimed from an atomic counter in increasing
//   tau = 6ks + 3(kend-k) + 2js + (j-jbeg)
// (every dependency has a smaller tau), so a claimed task only ever waits on
// tasks already claimed by running workgroups: the schedule cannot deadlock.
// Waits are bounded and raise an error flag instead of hanging.
//
// Line solver: the reference's pr/qr recurrence with every rounding of the
// reference (the SIMPLE iteration amplifies last-bit differences by ~1e7 per
// step, so the plain FP64 mode keeps every bit).  The pivots depend only on
// the coefficients and are computed for all lines up front (k_line_prep);
// the wavefront carries the right-hand-side recurrence, one wave per line,
// rows held in registers.

#include <algorithm>

#include "div_pos.h"
#include "fv.h"

namespace amcfd {

namespace {

constexpr int SW_MAXN = 1024;                 // longest x-line handled
constexpr int SW_M = 20;                      // rows per lane in the recurrence (64 * SW_M >= SW_MAXN)
constexpr int SW_M_SHORT = 6;                 // rows per lane for lines of at most 384 cells
constexpr int SW_M_THROUGHPUT = 8;            // rows per lane of the one-wave form (lines <= 512)
constexpr long long SPIN_LIMIT = 1LL << 20;   // bounded wait (~ a second)
static_assert(64 * SW_M >= SW_MAXN, "one wave must hold a whole line");

struct Coefs {
    const double *an, *as, *ae, *aw, *at, *ab, *ap, *su;
    const double *pr, *den, *y;   // per-line pivots from k_line_prep (exact path)
};

// "not written yet" marker of the pass buffers: a signalling NaN with a
// payload; arithmetic on the GPU only ever yields the canonical quiet NaN
constexpr unsigned long long SW_PENDING = 0x7ff4dead0000beefull;

__device__ __forceinline__ double lane_bcast(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}


// Per-line pivots of the reference recurrence.  den[i] = ap - aw*pr[i-1]
// (plus the melt-pool guard) and pr[i] = ae/den[i] depend only on the
// coefficients, which are fixed for the whole sweep, so they are computed
// once for every line of the sweep in parallel, with y[i] = RN(1/den[i]).
// One wave serves 64 lines (one lane each); the rows move through LDS in
// tiles of LP_CH so that global loads and stores are coalesced row segments,
// and the next tile is in flight while the lanes work through the current one.
constexpr int LP_CH = 16;

__global__ __launch_bounds__(64) void k_line_prep(Dims d, const double *__restrict__ aw,
                                                  const double *__restrict__ ap, const double *__restrict__ ae,
                                                  double *pr_o, double *den_o, double *y_o, int ibeg, int iend,
                                                  int jbeg, int jend, int kbeg, int kend, int guard) {
    constexpr int LD = LP_CH + 1;   // padded row stride of a line in the tiles
    __shared__ double T[3][64 * LD], O[3][64 * LD];
    const int lane = threadIdx.x;
    const int nj = jend - jbeg + 1;
    const int nlines = nj * (kend - kbeg + 1);
    const int line0 = blockIdx.x * 64;
    const int n = iend - ibeg + 1;
    const int nch = (n + LP_CH - 1) / LP_CH;
    // element e of a tile: line e / LP_CH of the wave, row e % LP_CH
    auto cell = [&](int e, int c, size_t &addr) -> bool {
        const int line = line0 + e / LP_CH, i = ibeg + c * LP_CH + e % LP_CH;
        if (line >= nlines || i > iend) return false;
        addr = (size_t)(kbeg + line / nj) * d.sk + (size_t)(jbeg + line % nj) * d.sj + i;
        return true;
    };
    double v0[LP_CH], v1[LP_CH], v2[LP_CH];
    auto fetch = [&](int c) {
#pragma unroll
        for (int u = 0; u < LP_CH; ++u) {
            size_t a;
            const bool ok = cell(lane + 64 * u, c, a);
            v0[u] = ok ? ap[a] : 1.0;
            v1[u] = ok ? aw[a] : 0.0;
            v2[u] = ok ? ae[a] : 0.0;
        }
    };
    fetch(0);
    double prv = 0.0;
    for (int c = 0; c < nch; ++c) {
#pragma unroll
        for (int u = 0; u < LP_CH; ++u) {
            const int e = lane + 64 * u;
            const int o = (e / LP_CH) * LD + e % LP_CH;
            T[0][o] = v0[u];
            T[1][o] = v1[u];
            T[2][o] = v2[u];
        }
        __syncthreads();
        if (c + 1 < nch) fetch(c + 1);
        const int rows = min(LP_CH, n - c * LP_CH);
        for (int r = 0; r < rows; ++r) {
            const int o = lane * LD + r;
            double den = T[0][o] - T[1][o] * prv;
            if (guard && fabs(den) <= 1e-12) den += AMCFD_SMALL;
            prv = T[2][o] / den;
            O[0][o] = prv;
            O[1][o] = den;
            O[2][o] = 1.0 / den;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < LP_CH; ++u) {
            const int e = lane + 64 * u;
            size_t a;
            if (cell(e, c, a)) {
                const int o = (e / LP_CH) * LD + e % LP_CH;
                pr_o[a] = O[0][o];
                den_o[a] = O[1][o];
                y_o[a] = O[2][o];
            }
        }
        __syncthreads();
    }
}

// Right-hand-side recurrence and back substitution of the reference
// (CFDSolverManager.cpp:2512-2521) for one line by one wave.  Lane l owns rows
// [l*M - pad, l*M - pad + M); rows are padded at the front to a multiple of M
// with pass-through rows (aw = 1, rhs = -0, den = y = 1, pr = 0: x = -0 + q
// is q for either sign of zero, and q / 1 is q), so there are no per-row
// bounds.  The running qr and the back-substitution value travel from lane to
// lane with readlane.  The coefficient registers are filled (load_coefs)
// before the line's dependencies are awaited; solve() then needs only the
// right-hand side.  LDS inputs (index = real row): AW, DEN, Y, PR, D = rhs;
// the solution overwrites D.
template <int M>
struct LineRegs {
    double raw[M], rden[M], ry[M];

    __device__ __forceinline__ void load_coefs(int n, const double *AW, const double *DEN, const double *Y) {
        const int lane = threadIdx.x & 63;
        const int nl = (n + M - 1) / M, pad = nl * M - n;
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int p = lane * M + r - pad;
            const bool real = (p >= 0) && (lane < nl);
            raw[r] = real ? AW[p] : 1.0;
            rden[r] = real ? DEN[p] : 1.0;
            ry[r] = real ? Y[p] : 1.0;
        }
    }

    __device__ __forceinline__ void solve(int n, const double *PR, double *D, double lo, double hi) {
        const int lane = threadIdx.x & 63;
        const int nl = (n + M - 1) / M, pad = nl * M - n;
        double rd[M], rx[M];
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int p = lane * M + r - pad;
            rd[r] = ((p >= 0) && (lane < nl)) ? D[p] : -0.0;
            rx[r] = 0.0;
        }
        double q = lo;
        for (int L = 0; L < nl; ++L) {
            if (lane == L) {
#pragma unroll
                for (int r = 0; r < M; ++r) {
                    const double x = rd[r] + raw[r] * q;
                    rx[r] = x;
                    q = div_pos(x, rden[r], ry[r]);
                    rd[r] = q;
                }
            }
            q = lane_bcast(q, L);
        }
        // the fast division's preconditions and its rounding, for every row
        // at once (no branches); a line that misses one is solved again with
        // IEEE division
        int bad = 0;
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const double ax = fabs(rx[r]);
            bad |= (int)!(rden[r] > 0.0) | (int)!(rden[r] < __builtin_inf()) |
                   ((int)(ax != 0.0) & ((int)!(ax >= 0x1p-900) | (int)!(ax < __builtin_inf()))) |
                   (int)!div_pos_exact(rx[r], rden[r], ry[r], rd[r]);
        }
        if (__any(bad)) {
            q = lo;
            for (int L = 0; L < nl; ++L) {
                if (lane == L) {
#pragma unroll
                    for (int r = 0; r < M; ++r) {
                        const int p = lane * M + r - pad;
                        const double dr = (p >= 0) ? D[p] : -0.0;
                        q = (dr + raw[r] * q) / rden[r];
                        rd[r] = q;
                    }
                }
                q = lane_bcast(q, L);
            }
        }
        double rpr[M];
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int p = lane * M + r - pad;
            rpr[r] = ((p >= 0) && (lane < nl)) ? PR[p] : 0.0;
        }
        double nxt = hi;
        for (int L = nl - 1; L >= 0; --L) {
            if (lane == L) {
#pragma unroll
                for (int r = M - 1; r >= 0; --r) {
                    nxt = rpr[r] * nxt + rd[r];
                    rd[r] = nxt;
                }
            }
            nxt = lane_bcast(nxt, L);
        }
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int p = lane * M + r - pad;
            if (p >= 0 && lane < nl) D[p] = rd[r];
        }
    }
};

// Fast-mode line solve.  With the pivots fixed for the sweep, both halves of
// the Thomas algorithm are first-order linear recurrences,
//   forward   q_i   = (aw_i y_i) q_{i-1} + d_i y_i,   q_{-1} = lo,
//   backward  phi_i = pr_i phi_{i+1} + q_i,           phi_n  = hi,
// so each is an inclusive scan of affine maps x -> a x + b: every lane
// composes the maps of its M consecutive rows, the lanes combine their maps
// with log2(64) shuffle steps, and each lane then replays its rows from the
// incoming value.  The rounding differs from the sequential recurrence.
// Rows past n are identity maps.
template <int M>
struct ScanLine {
    __device__ __forceinline__ static double shfl_up(double v, int off) { return __shfl_up(v, off, 64); }
    __device__ __forceinline__ static double shfl_down(double v, int off) { return __shfl_down(v, off, 64); }

    // D = rhs in, solution out; AW, Y, PR as in LineRegs
    __device__ static void solve(int n, const double *AW, const double *Y, const double *PR, double *D, double lo,
                                 double hi) {
        const int lane = threadIdx.x & 63;
        double a[M], b[M];
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int p = lane * M + r;
            const bool real = p < n;
            const double y = real ? Y[p] : 1.0;
            a[r] = real ? AW[p] * y : 1.0;
            b[r] = real ? D[p] * y : 0.0;
        }
        // forward: compose this lane's rows, scan over lanes, replay
        double A = 1.0, B = 0.0;
#pragma unroll
        for (int r = 0; r < M; ++r) {
            B = a[r] * B + b[r];
            A = a[r] * A;
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double Ap = shfl_up(A, off), Bp = shfl_up(B, off);
            if (lane >= off) {
                B = A * Bp + B;
                A = A * Ap;
            }
        }
        double Ae = shfl_up(A, 1), Be = shfl_up(B, 1);
        double x = (lane == 0) ? lo : Ae * lo + Be;
#pragma unroll
        for (int r = 0; r < M; ++r) {
            x = a[r] * x + b[r];
            b[r] = x;
        }
        // backward: the same from the top row down
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int p = lane * M + r;
            a[r] = (p < n) ? PR[p] : 1.0;
            if (p >= n) b[r] = 0.0;
        }
        A = 1.0;
        B = 0.0;
#pragma unroll
        for (int r = M - 1; r >= 0; --r) {
            B = a[r] * B + b[r];
            A = a[r] * A;
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double Ap = shfl_down(A, off), Bp = shfl_down(B, off);
            if (lane + off < 64) {
                B = A * Bp + B;
                A = A * Ap;
            }
        }
        Ae = shfl_down(A, 1);
        Be = shfl_down(B, 1);
        x = (lane == 63) ? hi : Ae * hi + Be;
#pragma unroll
        for (int r = M - 1; r >= 0; --r) {
            x = a[r] * x + b[r];
            b[r] = x;
        }
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int p = lane * M + r;
            if (p < n) D[p] = b[r];
        }
    }
};

// fewest rows per lane that fit the line
__device__ __forceinline__ void scan_solve(int n, const double *AW, const double *Y, const double *PR, double *D,
                                           double lo, double hi) {
    if (n <= 128) ScanLine<2>::solve(n, AW, Y, PR, D, lo, hi);
    else if (n <= 256) ScanLine<4>::solve(n, AW, Y, PR, D, lo, hi);
    else if (n <= 448) ScanLine<7>::solve(n, AW, Y, PR, D, lo, hi);
    else if (n <= 640) ScanLine<10>::solve(n, AW, Y, PR, D, lo, hi);
    else ScanLine<(SW_MAXN + 63) / 64>::solve(n, AW, Y, PR, D, lo, hi);
}

// Line data move between workgroups (and XCDs) with device-coherent
// relaxed atomic loads and stores (no L2-wide writeback or invalidate).  A
// consumer polls a pass-buffer value until it is no longer SW_PENDING.
__device__ __forceinline__ double ld_line(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_line(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool pending(double v) {
    return __builtin_bit_cast(unsigned long long, v) == SW_PENDING;
}
// Re-read v from p until it has been written; false if the bounded wait
// expired (or another wait of this launch already gave up).  Each read waits
// out its own memory round trip, so the loop needs no pause of its own (a
// pause of 64 or 128 cycles between reads made the sweeps slower).
__device__ bool await_value(const double *p, double &v, const int *error) {
    long long spins = 0;
    while (pending(v)) {
        v = ld_line(p);
        if (++spins > SPIN_LIMIT) return false;
        if ((spins & 255) == 0 && __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return false;
    }
    return true;
}

// pass-buffer fill before a sweep: every cell the sweep's tasks write
__global__ __launch_bounds__(256) void k_fill_pending(Dims d, double *pb, size_t pstride, int ibeg, int iend,
                                                      int jbeg, int jend, int kbeg, int kend) {
    const int n = iend - ibeg + 1, nj = jend - jbeg + 1;
    const size_t total = (size_t)n * nj * (kend - kbeg + 1);
    const double v = __builtin_bit_cast(double, SW_PENDING);
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const int i = ibeg + (int)(e % n);
        const size_t line = e / n;
        const int j = jbeg + (int)(line % nj), k = kbeg + (int)(line / nj);
        const size_t c = (size_t)k * d.sk + (size_t)j * d.sj + i;
#pragma unroll
        for (int p = 0; p < 4; ++p) pb[p * pstride + c] = v;
    }
}

// One x-line solve per task by a workgroup of NT threads; M rows per lane in
// the recurrence (lines up to LMAX = min(64 M, SW_MAXN) cells).  DOMAIN marks
// the instantiation for the whole-domain enthalpy sweep, so it shows up as
// its own kernel in a profile; the code is the same.  SCAN selects the
// fast-mode line solve (ScanLine) instead of the reference's recurrence.
// NT = 256 with M = 20 is the latency form: four waves gather the right-hand
// side, one runs the recurrence with few lane hand-offs, and the register
// file holds one such workgroup per CU.  NT = 64 with M = 8 is the throughput
// form for solvers sharing the device: one wave per line, eight per CU.  The
// recurrence performs the same operations in the same order either way.
template <int NT, int M, bool DOMAIN, bool SCAN>
__global__ __launch_bounds__(NT, (NT == 64 ? 2 : 1)) void k_line_sweeps(Dims d, Coefs cf, double *phi, double *pb, size_t pstride,
                                                    const int *__restrict__ tasks, int ntasks, int ibeg, int iend,
                                                    int jbeg, int jend, int kbeg, int kend,
                                                    unsigned long long *counter, unsigned long long base,
                                                    DevScalars *sc) {
    constexpr int LMAX = (64 * M < SW_MAXN) ? 64 * M : SW_MAXN;
    constexpr int MAXR = (LMAX + NT - 1) / NT;
    __shared__ double A[LMAX], B[LMAX], C[LMAX], D[LMAX], E[LMAX];
    __shared__ int s_task;
    const int tid = threadIdx.x;
    const int n = iend - ibeg + 1;
    LineRegs<M> regs;
    for (;;) {
        if (tid == 0) {
            unsigned long long t = atomicAdd(counter, 1ULL) - base;
            s_task = (t < (unsigned long long)ntasks) ? (int)t : -1;
        }
        __syncthreads();
        const int t = s_task;
        __syncthreads();
        if (t < 0) return;
        const int packed = tasks[t];
        const int pass = packed >> 28, k = (packed >> 14) & 0x3fff, j = packed & 0x3fff;
        const int ks = pass >> 1;
        const size_t row = (size_t)k * d.sk + (size_t)j * d.sj;
        // coefficients do not depend on other lines: stage them before waiting
        double rat[MAXR], rab[MAXR], ran[MAXR], ras[MAXR], rsu[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) {
                const size_t c = row + ibeg + ii;
                rat[r] = cf.at[c]; rab[r] = cf.ab[c]; ran[r] = cf.an[c]; ras[r] = cf.as[c]; rsu[r] = cf.su[c];
                A[ii] = cf.aw[c];
                if (!SCAN) B[ii] = cf.den[c];   // the scan solve needs only 1/den
                C[ii] = cf.y[c];
                E[ii] = cf.pr[c];
            }
        }
        const double lo = phi[row + ibeg - 1], hi = phi[row + iend + 1];   // fixed end values
        __syncthreads();
        if (!SCAN && tid < 64) regs.load_coefs(n, A, B, C);
        // neighbour lines as the sequential loop nest sees them: a pass buffer
        // inside the sweep's range, the sweep's input outside it
        const double *src_t = (k < kend) ? pb + (size_t)(2 * ks + 1) * pstride : phi;
        const double *src_b = (k > kbeg && ks == 1) ? pb + pstride : phi;
        const double *src_n = (j < jend && pass > 0) ? pb + (size_t)(pass - 1) * pstride : phi;
        const double *src_s = (j > jbeg) ? pb + (size_t)pass * pstride : phi;
        // right-hand side in the reference's order; every load is issued
        // first, then values still pending are polled
        double vt[MAXR], vb[MAXR], vn[MAXR], vs[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) {
                const size_t c = row + ibeg + ii;
                vt[r] = ld_line(src_t + c + d.sk);
                vb[r] = ld_line(src_b + c - d.sk);
                vn[r] = ld_line(src_n + c + d.sj);
                vs[r] = ld_line(src_s + c - d.sj);
            }
        }
        bool ok = true;
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) {
                const size_t c = row + ibeg + ii;
                ok &= await_value(src_t + c + d.sk, vt[r], &sc->error);
                ok &= await_value(src_b + c - d.sk, vb[r], &sc->error);
                ok &= await_value(src_n + c + d.sj, vn[r], &sc->error);
                ok &= await_value(src_s + c - d.sj, vs[r], &sc->error);
                D[ii] = rat[r] * vt[r] + rab[r] * vb[r] + ran[r] * vn[r] + ras[r] * vs[r] + rsu[r];
            }
        }
        if (!ok) atomicOr(&sc->error, 1);
        __syncthreads();
        if (tid < 64) {
            if constexpr (SCAN)
                scan_solve(n, A, C, E, D, lo, hi);
            else
                regs.solve(n, E, D, lo, hi);
        }
        __syncthreads();
        double *out = pb + (size_t)pass * pstride + row + ibeg;
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) {
                st_line(out + ii, D[ii]);
                if (pass == 3) phi[row + ibeg + ii] = D[ii];
            }
        }
    }
}

}  // namespace

void build_sweep_tasks(int kbeg, int kend, int jbeg, int jend, std::vector<int> &tasks, int &width) {
    tasks.clear();
    width = 0;
    if (kend < kbeg || jend < jbeg) return;
    const int nk = kend - kbeg + 1, nj = jend - jbeg + 1;
    const int tmax = 6 + 3 * (nk - 1) + 2 + (nj - 1);
    for (int tau = 0; tau <= tmax; ++tau) {
        int w = 0;
        for (int ks = 0; ks < 2; ++ks)
            for (int kk = 0; kk < nk; ++kk)
                for (int js = 0; js < 2; ++js) {
                    int jj = tau - 6 * ks - 3 * kk - 2 * js;
                    if (jj < 0 || jj >= nj) continue;
                    int k = kend - kk, j = jbeg + jj, pass = 2 * ks + js;
                    tasks.push_back((pass << 28) | (k << 14) | j);
                    ++w;
                }
        if (w > width) width = w;
    }
}

int sweep_max_line() { return SW_MAXN; }

void launch_line_prep(hipStream_t st, const Dims &d, DevFields f, double *pb, size_t pstride, const SweepPlan &plan,
                      int ibeg, int iend, int guard) {
    if (plan.ntasks == 0 || iend < ibeg) return;
    const int nlines = (plan.kend - plan.kbeg + 1) * (plan.jend - plan.jbeg + 1);
    hipLaunchKernelGGL(k_line_prep, dim3((nlines + 63) / 64), dim3(64), 0, st, d, f.aw, f.ap, f.ae, f.lpr, f.lden,
                       f.ly, ibeg, iend, plan.jbeg, plan.jend, plan.kbeg, plan.kend, guard);
    const size_t cells = (size_t)(iend - ibeg + 1) * nlines;
    const int fill_blocks = (int)std::min<size_t>((cells + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_pending, dim3(fill_blocks), dim3(256), 0, st, d, pb, pstride, ibeg, iend, plan.jbeg,
                       plan.jend, plan.kbeg, plan.kend);
}

template <bool SCAN>
static void launch_sweeps_t(hipStream_t st, const Dims &d, const Coefs &cf, double *phi, double *pb, size_t pstride,
                            const SweepPlan &plan, int ibeg, int iend, int guard, unsigned long long *counter,
                            unsigned long long counter_base, DevScalars *sc, int nblocks, int shared) {
    const int n = iend - ibeg + 1;
    // beside other solvers: one wave per line, for lines of up to 512 cells
    // (the scan mode keeps its lane layout, on which its rounding depends)
    if (shared && !SCAN && n <= 64 * SW_M_THROUGHPUT) {
        if (!guard)
            hipLaunchKernelGGL((k_line_sweeps<64, SW_M_THROUGHPUT, true, false>), dim3(nblocks), dim3(64), 0, st, d,
                               cf, phi, pb, pstride, plan.tasks, plan.ntasks, ibeg, iend, plan.jbeg, plan.jend,
                               plan.kbeg, plan.kend, counter, counter_base, sc);
        else
            hipLaunchKernelGGL((k_line_sweeps<64, SW_M_THROUGHPUT, false, false>), dim3(nblocks), dim3(64), 0, st, d,
                               cf, phi, pb, pstride, plan.tasks, plan.ntasks, ibeg, iend, plan.jbeg, plan.jend,
                               plan.kbeg, plan.kend, counter, counter_base, sc);
        return;
    }
    // short lines (the melt-pool box) get a lighter instantiation: fewer
    // registers, more waves per SIMD
    if (!guard)
        hipLaunchKernelGGL((k_line_sweeps<256, SW_M, true, SCAN>), dim3(nblocks), dim3(256), 0, st, d, cf, phi, pb,
                           pstride, plan.tasks, plan.ntasks, ibeg, iend, plan.jbeg, plan.jend, plan.kbeg, plan.kend,
                           counter, counter_base, sc);
    else if (n <= 64 * SW_M_SHORT)
        hipLaunchKernelGGL((k_line_sweeps<256, SW_M_SHORT, false, SCAN>), dim3(nblocks), dim3(256), 0, st, d, cf, phi,
                           pb, pstride, plan.tasks, plan.ntasks, ibeg, iend, plan.jbeg, plan.jend, plan.kbeg, plan.kend,
                           counter, counter_base, sc);
    else
        hipLaunchKernelGGL((k_line_sweeps<256, SW_M, false, SCAN>), dim3(nblocks), dim3(256), 0, st, d, cf, phi, pb,
                           pstride, plan.tasks, plan.ntasks, ibeg, iend, plan.jbeg, plan.jend, plan.kbeg, plan.kend,
                           counter, counter_base, sc);
}

void launch_line_sweeps(hipStream_t st, const Dims &d, DevFields f, double *phi, double *pb, size_t pstride,
                        const SweepPlan &plan, int ibeg, int iend, int guard, int scan, unsigned long long *counter,
                        unsigned long long counter_base, DevScalars *sc, int nblocks, int shared) {
    if (plan.ntasks == 0 || iend < ibeg) return;
    Coefs cf{f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su, f.lpr, f.lden, f.ly};
    if (scan)
        launch_sweeps_t<true>(st, d, cf, phi, pb, pstride, plan, ibeg, iend, guard, counter, counter_base, sc,
                              nblocks, shared);
    else
        launch_sweeps_t<false>(st, d, cf, phi, pb, pstride, plan, ibeg, iend, guard, counter, counter_base, sc,
                               nblocks, shared);
}

}  // namespace amcfd
