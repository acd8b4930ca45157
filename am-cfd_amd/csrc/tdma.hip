// Batched x-line TDMA on caller-owned device buffers.
//
// The reference solves each line with the recurrence of
// CFDSolverManager.cpp:2421-2450 (melt-pool box, with the |den| <= 1e-12
// guard) and 2500-2522 (whole domain):
//   pr[0] = 0, qr[0] = phi[0]
//   for i = 1 .. n-2:  den = ap - aw*pr[i-1]  (guard: den += 1e-6 if |den| <= 1e-12)
//                      pr[i] = ae/den,  qr[i] = (d + aw*qr[i-1])/den
//   for i = n-2 .. 1:  phi[i] = pr[i]*phi[i+1] + qr[i]
// where d is the line's right-hand side (su plus the neighbour-line terms).
// Here every line is one lane, in the same operations and order (compiled
// without contraction, so the result is the reference's bit for bit).  The
// pr/qr arrays are the scratch: amcfd_tdma_lines keeps a per-device
// workspace that it allocates and grows; amcfd_tdma_lines_noalloc uses a
// fixed file-scope device array instead and allocates nothing (the pr/qr
// entries are written before they are read, so nothing needs clearing).

#include <mutex>

#include "fv.h"

namespace amcfd {

namespace {

// Scratch layout: pr of row i of line l at [i*nlines + l], qr after all of
// pr, so that the lanes of a wave touch consecutive addresses.
__global__ __launch_bounds__(256) void k_tdma_lines(const double *__restrict__ aw, const double *__restrict__ ap,
                                                     const double *__restrict__ ae, const double *__restrict__ dv,
                                                     double *phi, int n, int nlines, int guard, double *scratch) {
    const int l = blockIdx.x * 256 + threadIdx.x;
    if (l >= nlines) return;
    const size_t base = (size_t)l * n;
    double *pr = scratch, *qr = scratch + (size_t)n * nlines;
    double p = 0.0, q = phi[base];
    pr[l] = p;
    qr[l] = q;
    for (int i = 1; i <= n - 2; ++i) {
        const size_t c = base + i;
        const double w = aw[c];
        double den = ap[c] - w * p;
        if (guard && fabs(den) <= 1e-12) den += AMCFD_SMALL;
        p = ae[c] / den;
        q = (dv[c] + w * q) / den;
        pr[(size_t)i * nlines + l] = p;
        qr[(size_t)i * nlines + l] = q;
    }
    double nxt = phi[base + n - 1];
    for (int i = n - 2; i >= 1; --i) {
        nxt = pr[(size_t)i * nlines + l] * nxt + qr[(size_t)i * nlines + l];
        phi[base + i] = nxt;
    }
}

int check_args(const double *aw, const double *ap, const double *ae, const double *d, double *phi, int n,
               int nlines) {
    if (!aw || !ap || !ae || !d || !phi) return set_error(AMCFD_E_ARG, "null buffer");
    if (n < 3) return set_error(AMCFD_E_ARG, "a line needs at least three nodes (two fixed ends)");
    if (nlines < 1) return set_error(AMCFD_E_ARG, "nlines must be positive");
    if ((long long)n * nlines > (1LL << 40)) return set_error(AMCFD_E_CAPACITY, "line batch too large");
    return AMCFD_OK;
}

int launch(const double *aw, const double *ap, const double *ae, const double *d, double *phi, int n, int nlines,
           int guard, double *scratch, hipStream_t st) {
    hipLaunchKernelGGL(k_tdma_lines, dim3((nlines + 255) / 256), dim3(256), 0, st, aw, ap, ae, d, phi, n, nlines,
                       guard, scratch);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(AMCFD_E_HIP, std::string("k_tdma_lines: ") + hipGetErrorString(e));
    return AMCFD_OK;
}

constexpr int MAX_DEV = 64;

// amcfd_tdma_lines: per-device workspace, allocated on first use and grown
struct Workspace {
    double *ptr = nullptr;
    size_t cap = 0;   // doubles
};
std::mutex ws_mutex;
Workspace ws[MAX_DEV];

// amcfd_tdma_lines_noalloc: fixed scratch, one copy per device
constexpr size_t NOALLOC_DOUBLES = (size_t)1 << 21;   // 16 MiB: 2*n*nlines <= 2^21
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];
std::mutex sym_mutex;
double *scratch_ptr[MAX_DEV] = {};

}  // namespace

}  // namespace amcfd

extern "C" {

int amcfd_tdma_lines(const double *aw, const double *ap, const double *ae, const double *d, double *phi, int n,
                     int nlines, int guard, void *stream) {
    using namespace amcfd;
    int st = check_args(aw, ap, ae, d, phi, n, nlines);
    if (st) return st;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV)
        return set_error(AMCFD_E_HIP, "no current HIP device");
    const size_t need = (size_t)2 * n * nlines;
    double *buf;
    {
        std::lock_guard<std::mutex> lock(ws_mutex);
        Workspace &w = ws[dev];
        if (w.cap < need) {
            // the old workspace may still be in use by queued work
            if (w.ptr) {
                if (hipDeviceSynchronize() != hipSuccess || hipFree(w.ptr) != hipSuccess)
                    return set_error(AMCFD_E_HIP, "workspace release failed");
                w.ptr = nullptr;
                w.cap = 0;
            }
            void *p = nullptr;
            if (hipMalloc(&p, need * sizeof(double)) != hipSuccess)
                return set_error(AMCFD_E_CAPACITY, "cannot allocate the TDMA workspace");
            w.ptr = (double *)p;
            w.cap = need;
        }
        buf = w.ptr;
    }
    return launch(aw, ap, ae, d, phi, n, nlines, guard, buf, (hipStream_t)stream);
}

int amcfd_tdma_lines_noalloc(const double *aw, const double *ap, const double *ae, const double *d, double *phi,
                             int n, int nlines, int guard, void *stream) {
    using namespace amcfd;
    int st = check_args(aw, ap, ae, d, phi, n, nlines);
    if (st) return st;
    if ((size_t)2 * n * nlines > NOALLOC_DOUBLES)
        return set_error(AMCFD_E_CAPACITY, "amcfd_tdma_lines_noalloc holds at most 2^20 line nodes (n * nlines)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV)
        return set_error(AMCFD_E_HIP, "no current HIP device");
    double *buf;
    {
        // the array's device address, looked up once per device with that
        // device current (the host-side name is not a device address)
        std::lock_guard<std::mutex> lock(sym_mutex);
        if (!scratch_ptr[dev]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return set_error(AMCFD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            scratch_ptr[dev] = (double *)p;
        }
        buf = scratch_ptr[dev];
    }
    return launch(aw, ap, ae, d, phi, n, nlines, guard, buf, (hipStream_t)stream);
}

}  // extern "C"
