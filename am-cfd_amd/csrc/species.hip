// Species conservation (ivar = 6) and the solidification parameters G and R.
//
//   applyBCs(6)                  BoundCondManager.cpp:655-699
//   discretizeEquations(6)       CFDSolverManager.cpp:1265-1367
//   calculateSource(6)           CFDSolverManager.cpp:1844-1924
//   calculateResiduals(6)        CFDSolverManager.cpp:2102-2130
//   enhanceConvergenceSpeed(6) and solveEntireDomain(6) reuse the enthalpy
//   kernels with the concentration as the unknown (energy.hip, sweep.hip)
//   updateSolutions, species part CFDSolverManager.cpp:2680-2685
//   getSolidificationParameters  CFDSolverManager.cpp:2763-2804
//
// Same layout and arithmetic conventions as the energy kernels: the
// reference's (k*nj + j)*ni + i arrays, each coefficient computed by the cell
// that owns it with the reference's operand order, residual partials folded
// by k_finalize.

#include "fv.h"

namespace amcfd {

namespace {

constexpr int TX = 64, TY = 4, KB = 4;

__device__ __forceinline__ size_t cix(const Dims &d, int i, int j, int k) {
    return (size_t)k * d.sk + (size_t)j * d.sj + (size_t)i;
}

}  // namespace

// ------------------------------------------------------------------------
// boundary values, three ordered phases as in the reference (the corners
// take the last copy): top/bottom of the interior columns, then whole
// north/south rows, then whole east/west columns
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bc_species_tb(Dims d, double tliquid, double topfac, double init,
                                                       const double *__restrict__ dzpbinv,
                                                       const double *__restrict__ temp, double *c) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.nim2 * d.njm2) return;
    const int i = 1 + t % d.nim2, j = 1 + t / d.nim2;
    const size_t top = cix(d, i, j, d.nkm1), below = cix(d, i, j, d.nkm2);
    const double cb = c[below];
    // 2*(13 - c)/dzpbinv/(dens*massdiff) + c where the surface is liquid
    c[top] = (temp[top] > tliquid) ? 2.0 * (13.0 - cb) / dzpbinv[d.nkm1] / topfac + cb : cb;
    c[cix(d, i, j, 0)] = init;
}

__global__ __launch_bounds__(256) void k_bc_species_ns(Dims d, double *c) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.ni * d.nk) return;
    const int i = t % d.ni, k = t / d.ni;
    c[cix(d, i, 0, k)] = c[cix(d, i, 1, k)];
    c[cix(d, i, d.njm1, k)] = c[cix(d, i, d.njm1 - 1, k)];
}

__global__ __launch_bounds__(256) void k_bc_species_ew(Dims d, double *c) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.nj * d.nk) return;
    const int j = t % d.nj, k = t / d.nj;
    c[cix(d, 0, j, k)] = c[cix(d, 1, j, k)];
    c[cix(d, d.nim1, j, k)] = c[cix(d, d.nim1 - 1, j, k)];
}

// ------------------------------------------------------------------------
// coefficients + boundary folding + relaxation + residual + block-correction
// terms.  The reference gives every face the diffusivity of the cell that
// writes it (1e-11 in solid, massdiff*rho in liquid; no harmonic mean) and
// twice the face velocity: a cell writes its north/east/top coefficient and
// its neighbour's south/west/bottom one, and the first row, column and plane
// get their south/west/bottom coefficient from the cell itself.  Gathered
// here: the south coefficient of row j >= 2 uses the diffusivity of row j-1.
// ------------------------------------------------------------------------
struct SpeciesConst {
    double tsolid, massdiff, delt, urfc;
    int scheme;
};

__device__ __forceinline__ double species_diff(double t, double r, double tsolid, double massdiff) {
    return (t <= tsolid) ? 1e-11 : massdiff * r;
}

__global__ __launch_bounds__(TX *TY) void k_species_assemble(
    Dims d, SpeciesConst q, DevGeom g, const double *__restrict__ uvel, const double *__restrict__ vvel,
    const double *__restrict__ wvel, const double *__restrict__ rho, const double *__restrict__ temp,
    const double *__restrict__ cnot, const double *__restrict__ cc, double *an_o, double *as_o, double *ae_o,
    double *aw_o, double *at_o, double *ab_o, double *ap_o, double *su_o, double *part_res, double *term_bl,
    double *term_blc) {
    __shared__ double scratch[2 * 4];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = 1 + blockIdx.x * TX + tx;
    const int j = 1 + blockIdx.y * TY + ty;
    const bool col_ok = (i <= d.nim2) && (j <= d.njm2);
    double sumd = 0.0, sumc = 0.0;
    const int sch = q.scheme;
    if (col_ok) {
        const double dxc = g.xu[i + 1] - g.xu[i], dyc = g.yv[j + 1] - g.yv[j];
        const double aij = g.areaij[j * d.ni + i];
        for (int kk = 0; kk < KB; ++kk) {
            const int k = 1 + blockIdx.z * KB + kk;
            if (k > d.nkm2) break;
            const size_t c = cix(d, i, j, k);
            const double aik = g.areaik[k * d.ni + i], ajk = g.areajk[k * d.nj + j];
            const double rP = rho[c];
            const double dP = species_diff(temp[c], rP, q.tsolid, q.massdiff);
            double an, as, ae, aw, at, ab;
            {   // north: this cell's face
                double vn = vvel[c + d.sj] * 2.0;
                double fn = rho[c + d.sj] * vn * aik;
                an = coef_minus(sch, dP * aik * g.dypsinv[j + 1], fn);
            }
            {   // south: written by the cell below (its north face), or by this cell in row 1
                double vs = vvel[c] * 2.0;
                double fs = rP * vs * aik;
                double dd = (j > 1) ? species_diff(temp[c - d.sj], rho[c - d.sj], q.tsolid, q.massdiff) * aik *
                                          g.dypsinv[j]
                                    : dP * aik * g.dypsinv[1];
                as = coef_plus(sch, dd, fs);
            }
            {   // east
                double ue = uvel[c + 1] * 2.0;
                double fe = rho[c + 1] * ue * ajk;
                ae = coef_minus(sch, dP * ajk * g.dxpwinv[i + 1], fe);
            }
            {   // west
                double uw = uvel[c] * 2.0;
                double fw = rP * uw * ajk;
                double dd = (i > 1) ? species_diff(temp[c - 1], rho[c - 1], q.tsolid, q.massdiff) * ajk *
                                          g.dxpwinv[i]
                                    : dP * ajk * g.dxpwinv[1];
                aw = coef_plus(sch, dd, fw);
            }
            {   // top
                double wt = wvel[c + d.sk] * 2.0;
                double ft = rho[c + d.sk] * wt * aij;
                at = coef_minus(sch, dP * aij * g.dzpbinv[k + 1], ft);
            }
            {   // bottom
                double wb = wvel[c] * 2.0;
                double fb = rP * wb * aij;
                double dd = (k > 1) ? species_diff(temp[c - d.sk], rho[c - d.sk], q.tsolid, q.massdiff) * aij *
                                          g.dzpbinv[k]
                                    : dP * aij * g.dzpbinv[1];
                ab = coef_plus(sch, dd, fb);
            }
            const double vol = dxc * dyc * (g.zw[k + 1] - g.zw[k]);
            const double apnot = rP / q.delt * vol;
            double sp = 0.0;
            double su = apnot * cnot[c];
            const double cP = cc[c];
            const double cN = cc[c + d.sj], cS = cc[c - d.sj], cE = cc[c + 1], cW = cc[c - 1];
            const double cT = cc[c + d.sk], cB = cc[c - d.sk];
            // boundary values folded into the source (bottom, top, south, north, west, east)
            if (k == 1) { su += ab * cB; sp -= ab; ab = 0.0; }
            if (k == d.nkm2) { su += at * cT; sp -= at; at = 0.0; }
            if (j == 1) { su += as * cS; sp -= as; as = 0.0; }
            if (j == d.njm2) { su += an * cN; sp -= an; an = 0.0; }
            if (i == 1) { su += aw * cW; sp -= aw; aw = 0.0; }
            if (i == d.nim2) { su += ae * cE; sp -= ae; ae = 0.0; }
            double ap = an + as + ae + aw + at + ab + apnot - sp;
            ap /= q.urfc;
            su += (1.0 - q.urfc) * ap * cP;
            an_o[c] = an; as_o[c] = as; ae_o[c] = ae; aw_o[c] = aw; at_o[c] = at; ab_o[c] = ab;
            ap_o[c] = ap; su_o[c] = su;
            double r = (an * cN + as * cS + ae * cE + aw * cW + at * cT + ab * cB + su) / ap - cP;
            sumd += fabs(r);
            sumc += fabs(cP);
            term_bl[c] = ap - an - as - at - ab;
            term_blc[c] = ae * cE + aw * cW + an * cN + as * cS + at * cT + ab * cB + su - ap * cP;
        }
    }
    double v[2] = {sumd, sumc};
    block_sum<2>(v, scratch);   // contains __syncthreads
    if (tx == 0 && ty == 0) {
        const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        part_res[2 * b] = v[0];
        part_res[2 * b + 1] = v[1];
    }
}

// end of step: c_old = c, liquid-side concentration c/(fl + kp(1-fl))
__global__ __launch_bounds__(256) void k_species_end(size_t n, double kp, const double *__restrict__ fracl,
                                                     const double *__restrict__ cc, double *cnot, double *cl) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double c = cc[t], fl = fracl[t];
    cnot[t] = c;
    cl[t] = c / (fl + kp * (1.0 - fl));
}

// ------------------------------------------------------------------------
// G and R: in every x-row (k < nkm1, all j), at the first node i with
// T >= Ts, the gradient and growth rate are written to node i-1 from the
// temperatures at flat offsets -1 / +1 (i-2, i), +-ni (rows) and +-ni*nj
// (planes) of node i-1, exactly as the reference indexes its flat array (a
// row at the edge reads into the adjacent row or plane).  Where the
// reference would read outside its coordinate or temperature arrays (first
// node at i = 0, or row j = 0, or plane k = 0) G and R become NaN.
// One wave per row; a ballot finds the first node.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_solidification(Dims d, double tsolid, double scanvel,
                                                        const double *__restrict__ x, const double *__restrict__ y,
                                                        const double *__restrict__ z, const double *__restrict__ temp,
                                                        double *grad, double *rate) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= d.nkm1 * d.nj) return;   // whole waves leave together
    const int j = row % d.nj, k = row / d.nj;
    const size_t base = (size_t)k * d.sk + (size_t)j * d.sj;
    int first = -1;
    for (int i0 = 0; i0 < d.ni && first < 0; i0 += 64) {
        const int i = i0 + lane;
        const bool hot = (i < d.ni) && (temp[base + i] >= tsolid);
        const unsigned long long m = __ballot(hot);
        if (m) first = i0 + __ffsll((long long)m) - 1;
    }
    if (first < 0 || lane != 0) return;
    const int i = first;
    const long long c = (long long)base + i, o = c - 1;
    if (i == 0 || j == 0 || k == 0) {
        if (o >= 0) {
            grad[o] = __longlong_as_double(0x7ff8000000000000LL);
            rate[o] = __longlong_as_double(0x7ff8000000000000LL);
        }
        return;
    }
    const long long sj = (long long)d.sj, sk = (long long)d.sk;
    const double gx = (temp[c] - temp[c - 2]) / (x[i] - x[i - 1]) / 2;
    const double gy = (temp[o + sj] - temp[o - sj]) / (y[j] - y[j - 1]) / 2;
    const double gz = (temp[o + sk] - temp[o - sk]) / (z[k] - z[k - 1]) / 2;
    const double gm = sqrt(gx * gx + gy * gy + gz * gz);
    grad[o] = gm;
    rate[o] = scanvel * gx / gm;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static dim3 interior_grid(const Dims &d) {
    return dim3((d.nim2 + TX - 1) / TX, (d.njm2 + TY - 1) / TY, (d.nkm2 + KB - 1) / KB);
}

void launch_bc_species(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, const amcfd_params &p) {
    const int n1 = d.nim2 * d.njm2, n2 = d.ni * d.nk, n3 = d.nj * d.nk;
    hipLaunchKernelGGL(k_bc_species_tb, dim3((n1 + 255) / 256), dim3(256), 0, st, d, p.tliquid, p.dens * p.massdiff,
                       p.initconcentration, g.dzpbinv, f.temp, f.conc);
    hipLaunchKernelGGL(k_bc_species_ns, dim3((n2 + 255) / 256), dim3(256), 0, st, d, f.conc);
    hipLaunchKernelGGL(k_bc_species_ew, dim3((n3 + 255) / 256), dim3(256), 0, st, d, f.conc);
}

void launch_species_assemble(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                             const amcfd_params &p) {
    SpeciesConst q{p.tsolid, p.massdiff, p.delt, p.urfc, p.discretization};
    hipLaunchKernelGGL(k_species_assemble, interior_grid(d), dim3(TX, TY), 0, st, d, q, g, f.uvel, f.vvel, f.wvel,
                       f.rho, f.temp, f.concnot, f.conc, f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su, pr.species,
                       f.sp, f.apnot);
}

void launch_species_end(hipStream_t st, const Dims &d, DevFields f, const amcfd_params &p) {
    const size_t n = (size_t)d.ni * d.nj * d.nk;
    hipLaunchKernelGGL(k_species_end, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p.kp, f.fracl, f.conc,
                       f.concnot, f.concl);
}

void launch_solidification(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, double tsolid,
                           double scanvel) {
    const int rows = d.nkm1 * d.nj;
    hipLaunchKernelGGL(k_solidification, dim3((rows + 3) / 4), dim3(256), 0, st, d, tsolid, scanvel, g.x, g.y, g.z,
                       f.temp, f.grad, f.rate);
}

}  // namespace amcfd
