// Moving mesh: in-place remap of the solution onto the regenerated grid,
// CFDSolverManager::remapSolutions (CFDSolverManager.cpp:2809-3051) with
// CFDSolverManager::interpolate (:3057-3066).
//
// The reference walks k, j (< nj-1), i (1..ni-2) over the active extent and,
// where a node moved by more than 1e-6, replaces each of twelve fields
// (fourteen with species) by the linear interpolant between the node and its
// neighbour in the direction of the move, evaluated at the node's new
// coordinate: first the whole x pass, then the whole y pass.  Walking upwards in place, a backward move
// (dir = -1) reads the neighbour that was already remapped, so each x-line
// (x pass) or y-column (y pass) is a recurrence; a forward move reads only
// values not yet touched.  One thread here owns one line or column of one
// field and walks it in the reference's order, so both directions give the
// reference's values bit for bit.  The y pass is coalesced (neighbouring
// threads take neighbouring i); the x pass strides by a line, which costs
// little next to a time step (twelve fields, once per step).

#include "internal.h"

namespace amcfd {

// the remapped fields, in the reference's order (the last two with species)
struct RemapFields {
    double *f[14];
};

__device__ __forceinline__ double lerp_ref(double x0, double x1, double y0, double y1, double x) {
    const double dy = y1 - y0, dx = x1 - x0;
    return y0 + (x - x0) * (dy / dx);
}

// one thread per (field, x-line j<nj-1, k<nk): walk i = 1..ni-2
__global__ void k_remap_x(Dims d, RemapFields rf, const double *__restrict__ xnew, const double *__restrict__ xold,
                          int dir) {
    const int lines = (d.nj - 1) * d.nk;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lines) return;
    const int j = t % (d.nj - 1), k = t / (d.nj - 1);
    double *row = rf.f[blockIdx.y] + (size_t)k * d.sk + (size_t)j * d.sj;
    for (int i = 1; i < d.ni - 1; ++i) {
        const double x0 = xold[i], xn = xnew[i];
        if (!(fabs(x0 - xn) > 1e-6)) continue;
        row[i] = lerp_ref(x0, xold[i + dir], row[i], row[i + dir], xn);
    }
}

// one thread per (field, i in 1..ni-2, k<nk): walk j = 0..nj-2
__global__ void k_remap_y(Dims d, RemapFields rf, const double *__restrict__ ynew, const double *__restrict__ yold,
                          int dir) {
    const int cols = (d.ni - 2) * d.nk;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cols) return;
    const int i = 1 + t % (d.ni - 2), k = t / (d.ni - 2);
    double *col = rf.f[blockIdx.y] + (size_t)k * d.sk + i;
    const long step = (long)dir * (long)d.sj;
    for (int j = 0; j < d.nj - 1; ++j) {
        const double y0 = yold[j], yn = ynew[j];
        // y[0] is 0 on every grid, so j = 0 never moves and dir = -1 never
        // reaches below the column; the second test only keeps it that way
        if (!(fabs(y0 - yn) > 1e-6) || j + dir < 0) continue;
        double *c = col + (size_t)j * d.sj;
        c[0] = lerp_ref(y0, yold[j + dir], c[0], c[step], yn);
    }
}

void launch_remap(hipStream_t st, const Dims &d, DevFields f, const DevGeom &g, const double *xold,
                  const double *yold, int dirx, int diry, int species) {
    RemapFields rf{{f.enthalpy, f.hnot, f.temp, f.fracl, f.uvel, f.vvel, f.wvel, f.unot, f.vnot, f.wnot,
                    f.pressure, f.pp, f.conc, f.concnot}};
    const unsigned nf = species ? 14 : 12;
    const int nt = 128;
    const int lines = (d.nj - 1) * d.nk, cols = (d.ni - 2) * d.nk;
    hipLaunchKernelGGL(k_remap_x, dim3((lines + nt - 1) / nt, nf), dim3(nt), 0, st, d, rf, g.x, xold, dirx);
    hipLaunchKernelGGL(k_remap_y, dim3((cols + nt - 1) / nt, nf), dim3(nt), 0, st, d, rf, g.y, yold, diry);
}

}  // namespace amcfd
