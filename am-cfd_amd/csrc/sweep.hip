// Gauss-Seidel-ordered x-line sweeps on the GPU.
//
// The reference sweeps  for ks in {0,1}: for k = kend..kbeg: for js in {0,1}:
// for j = jbeg..jend: Thomas along i  (CFDSolverManager.cpp:2415-2453 on the
// melt-pool box, 2492-2525 on the whole domain).  Each line's right-hand side
// reads the lines (k+-1, j) and (k, j+-1) as they stand at that point of the
// sequence, so the sweep is Gauss-Seidel in j and k.  Run single-threaded
// that order is the reference's definition of the result.
//
// Here every line solve is a task; task (pass p = 2ks+js, k, j) may start once
//   line (k, j-1) has finished pass p,
//   line (k, j+1) has finished pass p-1            (p > 0),
//   line (k+1, j) has finished pass 2ks+1,
//   line (k-1, j) has finished pass 1              (ks = 1),
// which is exactly the set of read-after-write and write-after-read
// constraints of the sequential loop nest.  Tasks are claimed from an atomic
// counter in increasing  tau = 6ks + 3(kend-k) + 2js + (j-jbeg)  (every
// dependency has a smaller tau), so a claimed task only ever waits on tasks
// already claimed by running workgroups: the schedule cannot deadlock.  Each
// line keeps a version word (epoch + passes done); waits are bounded and raise
// an error flag instead of hanging.
//
// Line solvers:
//   EXACT  the reference's pr/qr recurrence, run by one lane from LDS after
//          the wave has staged the line; bitwise the reference's arithmetic
//          (the SIMPLE iteration amplifies last-bit differences by ~1e7 per
//          step, so the plain FP64 mode keeps every bit).
//   PCR    parallel cyclic reduction by the whole workgroup; same system,
//          different rounding (for modes that do not promise bitwise
//          agreement).  On the melt-pool box the reference guards
//          |pivot| <= 1e-12; a line that is not provably free of such pivots
//          falls back to the exact recurrence.

#include "fv.h"

namespace amcfd {

namespace {

constexpr int SW_MAXN = 1024;                 // longest x-line handled
constexpr long long SPIN_LIMIT = 1LL << 20;   // bounded wait (~ a second)

struct Coefs {
    const double *an, *as, *ae, *aw, *at, *ab, *ap, *su;
    const double *pr, *den, *y;   // per-line pivots from k_line_prep (exact path)
};

__device__ __forceinline__ int load_version(const int *v) {
    return __hip_atomic_load(v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until *v >= need.  Gives up after SPIN_LIMIT polls, or at once when
// another wait of this launch has already given up (error flag set).
__device__ bool wait_version(const int *v, int need, const int *error) {
    long long spins = 0;
    while (load_version(v) < need) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > SPIN_LIMIT) return false;
        if ((spins & 255) == 0 && load_version(error) != 0) return false;
    }
    return true;
}

// Reference recurrence on LDS rows (A = -aw, B = ap, C = -ae, D = rhs);
// the solution overwrites D.  Run by a single lane.
__device__ void thomas_exact(int n, double *A, const double *B, double *C, double *D, double lo, double hi,
                             int guard) {
    double prv = 0.0, qrv = lo;
    for (int ii = 0; ii < n; ++ii) {
        const double aw = -A[ii];
        double den = B[ii] - aw * prv;
        if (guard && fabs(den) <= 1e-12) den += AMCFD_SMALL;
        prv = -C[ii] / den;
        qrv = (D[ii] + aw * qrv) / den;
        A[ii] = prv;
        C[ii] = qrv;
    }
    double nxt = hi;
    for (int ii = n - 1; ii >= 0; --ii) {
        nxt = A[ii] * nxt + C[ii];
        D[ii] = nxt;
    }
}

__device__ __forceinline__ double lane_bcast(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Correctly rounded x / b given y = RN(1/b): a faithful first quotient and
// two exact-remainder corrections (Markstein's theorem: with y the rounded
// reciprocal and q faithful, RN(q + (x - b q) y) = RN(x / b)).  Five dependent
// operations instead of the eleven of a full division; valid for the normal,
// finite operands of the line solves (pivots ~1e-7 .. 1e20).  A zero x keeps
// the sign IEEE division gives it.
__device__ __forceinline__ double div_by_recip(double x, double b, double y) {
    const double q0 = x * y;
    const double e0 = __fma_rn(-b, q0, x);
    const double q1 = __fma_rn(e0, y, q0);
    const double e1 = __fma_rn(-b, q1, x);
    const double q2 = __fma_rn(e1, y, q1);
    return (x == 0.0) ? q0 : q2;
}

// Per-line pivots of the reference recurrence.  den[i] = ap - aw*pr[i-1]
// (plus the melt-pool guard) and pr[i] = ae/den[i] depend only on the
// coefficients, which are fixed for the whole sweep, so they are computed
// once for every line of the sweep in parallel (one lane per line), with
// y[i] = RN(1/den[i]).  The wavefront then only carries the right-hand-side
// recurrence.
__global__ __launch_bounds__(256) void k_line_prep(Dims d, const double *__restrict__ aw,
                                                   const double *__restrict__ ap, const double *__restrict__ ae,
                                                   double *pr_o, double *den_o, double *y_o, int ibeg, int iend,
                                                   int jbeg, int jend, int kbeg, int kend, int guard) {
    const int nj = jend - jbeg + 1;
    const int line = blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nj * (kend - kbeg + 1)) return;
    const int k = kbeg + line / nj, j = jbeg + line % nj;
    const size_t row = (size_t)k * d.sk + (size_t)j * d.sj;
    double prv = 0.0;
    for (int i = ibeg; i <= iend; ++i) {
        const size_t c = row + i;
        double den = ap[c] - aw[c] * prv;
        if (guard && fabs(den) <= 1e-12) den += AMCFD_SMALL;
        prv = ae[c] / den;
        pr_o[c] = prv;
        den_o[c] = den;
        y_o[c] = 1.0 / den;
    }
}

// Right-hand-side recurrence and back substitution of the reference
// (CFDSolverManager.cpp:2512-2521) for one line by one wave, rows in
// registers.  Rows are padded at the front to NL*M (pass-through rows:
// aw = 1, rhs = 0, den = y = 1, pr = 0) so lane l owns rows [l*M, l*M + M)
// with no per-row bounds; the running qr and the back-substitution value
// travel from lane to lane with readlane.  LDS inputs (index = real row):
// AW, DEN, Y, PR, D = rhs; the solution overwrites D.
template <int M>
__device__ void rhs_wave(int n, const double *AW, const double *DEN, const double *Y, const double *PR, double *D,
                         double lo, double hi) {
    const int lane = threadIdx.x;
    const int nl = (n + M - 1) / M;
    const int pad = nl * M - n;
    double raw[M], rden[M], ry[M], rpr[M], rd[M];
#pragma unroll
    for (int r = 0; r < M; ++r) {
        const int p = lane * M + r - pad;
        const bool real = (p >= 0) && (lane < nl);
        raw[r] = real ? AW[p] : 1.0;
        rden[r] = real ? DEN[p] : 1.0;
        ry[r] = real ? Y[p] : 1.0;
        rpr[r] = real ? PR[p] : 0.0;
        rd[r] = real ? D[p] : 0.0;
    }
    double qrv = lo;
    for (int L = 0; L < nl; ++L) {
        if (lane == L) {
#pragma unroll
            for (int r = 0; r < M; ++r) {
                qrv = div_by_recip(rd[r] + raw[r] * qrv, rden[r], ry[r]);
                rd[r] = qrv;
            }
        }
        qrv = lane_bcast(qrv, L);
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

template <int M>
__device__ void rhs_wave_dispatch(int n, const double *AW, const double *DEN, const double *Y, const double *PR,
                                  double *D, double lo, double hi) {
    if constexpr (M < 16) {
        if (n > 64 * M) {
            rhs_wave_dispatch<M + 1>(n, AW, DEN, Y, PR, D, lo, hi);
            return;
        }
    }
    rhs_wave<M>(n, AW, DEN, Y, PR, D, lo, hi);
}

// Line data move between workgroups (and XCDs) with device-coherent
// relaxed atomic loads and stores, so no L2-wide writeback / invalidate is
// needed: a producer stores its line, waits for the stores to complete
// (s_waitcnt) and only then publishes the line's version word; a consumer
// reads the neighbour lines only after it has seen their versions.
__device__ __forceinline__ double ld_line(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_line(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NT, bool PCR>
__global__ __launch_bounds__(NT) void k_line_sweeps(Dims d, Coefs cf, double *phi, const int *__restrict__ tasks,
                                                    int ntasks, int ibeg, int iend, int jbeg, int jend, int kbeg,
                                                    int kend, int guard, unsigned long long *counter,
                                                    unsigned long long base, int *versions, int epoch,
                                                    DevScalars *sc) {
    constexpr int MAXR = SW_MAXN / NT;
    __shared__ double A[SW_MAXN], B[SW_MAXN], C[SW_MAXN], D[SW_MAXN], E[PCR ? 1 : SW_MAXN];
    __shared__ int s_task, s_ok;
    const int tid = threadIdx.x;
    const int n = iend - ibeg + 1;
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
        double rat[MAXR], rab[MAXR], ran[MAXR], ras[MAXR], rsu[MAXR], raw[MAXR], rap[MAXR], rae[MAXR];
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) {
                const size_t c = row + ibeg + ii;
                rat[r] = cf.at[c]; rab[r] = cf.ab[c]; ran[r] = cf.an[c]; ras[r] = cf.as[c]; rsu[r] = cf.su[c];
                raw[r] = cf.aw[c];
                if (PCR) { rap[r] = cf.ap[c]; rae[r] = cf.ae[c]; }
                else { rap[r] = cf.den[c]; rae[r] = cf.y[c]; }
            }
        }
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) {
                if (PCR) {   // a x[i-1] + b x[i] + c x[i+1] = d with a = -aw, c = -ae
                    A[ii] = -raw[r];
                    B[ii] = rap[r];
                    C[ii] = -rae[r];
                } else {     // exact recurrence: aw, pivot, 1/pivot, pr
                    A[ii] = raw[r];
                    B[ii] = rap[r];
                    C[ii] = rae[r];
                    E[ii] = cf.pr[row + ibeg + ii];
                }
            }
        }
        const double lo = phi[row + ibeg - 1], hi = phi[row + iend + 1];   // fixed end values
        if (tid == 0) {
            bool ok = true;
            const int *err = &sc->error;
            const int *vb = versions;
            if (j > jbeg) ok &= wait_version(vb + k * d.nj + (j - 1), epoch + pass + 1, err);
            if (j < jend && pass > 0) ok &= wait_version(vb + k * d.nj + (j + 1), epoch + pass, err);
            if (k < kend) ok &= wait_version(vb + (k + 1) * d.nj + j, epoch + 2 * ks + 2, err);
            if (k > kbeg && ks > 0) ok &= wait_version(vb + (k - 1) * d.nj + j, epoch + 2, err);
            if (!ok) atomicOr(&sc->error, 1);
            s_ok = PCR ? 1 : 0;
        }
        __syncthreads();
        // right-hand side from the neighbouring lines, in the reference's order
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) {
                const size_t c = row + ibeg + ii;
                D[ii] = rat[r] * ld_line(phi + c + d.sk) + rab[r] * ld_line(phi + c - d.sk) +
                        ran[r] * ld_line(phi + c + d.sj) + ras[r] * ld_line(phi + c - d.sj) + rsu[r];
            }
        }
        if (PCR && guard) {
            // PCR only when no Thomas pivot can fall to 1e-12: nonnegative
            // neighbours, ap >= aw + ae and ap - aw > 2e-12
            bool ok = true;
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                const int ii = tid + r * NT;
                if (ii < n) {
                    const double aw = raw[r], ae = rae[r], ap = rap[r];
                    ok &= (aw >= 0.0) && (ae >= 0.0) && (ap - aw > 2e-12) && (ap >= (aw + ae) * (1.0 - 1e-12));
                }
            }
            if (!ok) s_ok = 0;
        }
        __syncthreads();
        if (PCR && s_ok) {
            if (tid == 0) {   // fixed end values fold into the first / last rows
                D[0] -= A[0] * lo;
                D[n - 1] -= C[n - 1] * hi;
                A[0] = 0.0;
                C[n - 1] = 0.0;
            }
            __syncthreads();
            pcr_solve<MAXR>(n, A, B, C, D);
        } else if (!PCR) {
            if (tid < 64) rhs_wave_dispatch<1>(n, A, B, C, E, D, lo, hi);
            __syncthreads();
        } else {
            if (tid == 0) thomas_exact(n, A, B, C, D, lo, hi, guard);
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < MAXR; ++r) {
            const int ii = tid + r * NT;
            if (ii < n) st_line(phi + row + ibeg + ii, D[ii]);
        }
        __builtin_amdgcn_s_waitcnt(0);   // this thread's line stores have completed
        __syncthreads();
        if (tid == 0)
            __hip_atomic_store(versions + k * d.nj + j, epoch + pass + 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
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

void launch_line_sweeps(hipStream_t st, const Dims &d, DevFields f, double *phi, const SweepPlan &plan, int ibeg,
                        int iend, int guard, unsigned long long *counter, unsigned long long counter_base,
                        int *versions, int epoch, DevScalars *sc, int nblocks) {
    if (plan.ntasks == 0 || iend < ibeg) return;
    Coefs cf{f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su, f.lpr, f.lden, f.ly};
    const int nlines = (plan.kend - plan.kbeg + 1) * (plan.jend - plan.jbeg + 1);
    hipLaunchKernelGGL(k_line_prep, dim3((nlines + 255) / 256), dim3(256), 0, st, d, f.aw, f.ap, f.ae, f.lpr, f.lden,
                       f.ly, ibeg, iend, plan.jbeg, plan.jend, plan.kbeg, plan.kend, guard);
    hipLaunchKernelGGL((k_line_sweeps<256, false>), dim3(nblocks), dim3(256), 0, st, d, cf, phi, plan.tasks,
                       plan.ntasks, ibeg, iend, plan.jbeg, plan.jend, plan.kbeg, plan.kend, guard, counter,
                       counter_base, versions, epoch, sc);
}

void launch_line_sweeps_pcr(hipStream_t st, const Dims &d, DevFields f, double *phi, const SweepPlan &plan,
                            int ibeg, int iend, int guard, unsigned long long *counter,
                            unsigned long long counter_base, int *versions, int epoch, DevScalars *sc,
                            int nblocks) {
    if (plan.ntasks == 0 || iend < ibeg) return;
    Coefs cf{f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL((k_line_sweeps<256, true>), dim3(nblocks), dim3(256), 0, st, d, cf, phi, plan.tasks,
                       plan.ntasks, ibeg, iend, plan.jbeg, plan.jend, plan.kbeg, plan.kend, guard, counter,
                       counter_base, versions, epoch, sc);
}

}  // namespace amcfd
