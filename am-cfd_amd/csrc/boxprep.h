// Readiness pattern and set-up of the persistent melt-pool box passes
// (mixed mode), shared by the box-pass kernels and the flow boundary kernel
// that prepares their buffers in the same launch.
#pragma once

#include "fv.h"

namespace amcfd {
namespace {

// a NaN no arithmetic produces: a pass-buffer value that is not written yet
constexpr unsigned long long BX_PENDING = 0x7ff4dead0000beefull;

struct BoxRange {
    int ibeg, iend, jbeg, jend, kbeg, kend;
};
// the box's line range from the device box (istat, iend, jstat, jend,
// kstat) and the top interior plane kend; false when empty or when a line
// is longer than maxn
__device__ __forceinline__ bool box_range(const int *dbox, int kend, int maxn, BoxRange &b) {
    b.ibeg = dbox[0] + 1;
    b.iend = dbox[1] - 1;
    b.jbeg = dbox[2];
    b.jend = dbox[3];
    b.kbeg = dbox[4];
    b.kend = kend;
    return !(b.iend - b.ibeg + 1 > maxn || b.iend < b.ibeg || b.jend < b.jbeg || b.kend < b.kbeg);
}

// the pass buffers' box cells to BX_PENDING and the task counter to zero
// (256 threads per workgroup)
struct BoxPrepK {
    static constexpr int family = 13, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, double *bufs, size_t bstride, int passes, int kend,
                                               unsigned int *claim, const int *__restrict__ skip,
                                               const int *__restrict__ dbox) {
        if (skip && *skip) return;
        BoxRange b;
        if (!box_range(dbox, kend, 1 << 30, b)) return;
        if (vb.x == 0 && threadIdx.x == 0) *claim = 0u;
        const int n = b.iend - b.ibeg + 1, nj = b.jend - b.jbeg + 1;
        const size_t cells = (size_t)n * nj * (b.kend - b.kbeg + 1);
        const double pend = __builtin_bit_cast(double, BX_PENDING);
        for (size_t e = (size_t)vb.x * 256 + threadIdx.x; e < cells * passes; e += (size_t)vb.gx * 256) {
            const size_t q = e / cells, r = e % cells;
            const int i = b.ibeg + (int)(r % n);
            const size_t line = r / n;
            const int j = b.jbeg + (int)(line % nj), k = b.kbeg + (int)(line / nj);
            bufs[q * bstride + (size_t)k * d.sk + (size_t)j * d.sj + (size_t)i] = pend;
        }
    }
};

}  // namespace
}  // namespace amcfd
