// Energy-equation kernels of one SIMPLE iteration (FP64 plain mode).
//
// Data layout: every field is the reference's flattened (k*nj + j)*ni + i
// array, i fastest, so each kernel reads x-contiguous rows; blocks are
// 64 (i) x 4 (j) threads marching over a few k planes.  Reductions write
// per-block partials that a single-workgroup kernel folds in a fixed order,
// so every run is bitwise reproducible.

#include "fv.h"

namespace amcfd {

namespace {

constexpr int TX = 64, TY = 4, KB = 4;
static_assert(TY == 4, "the mixed-mode column partials give one quantity to each of the four thread rows");

__device__ __forceinline__ size_t cix(const Dims &d, int i, int j, int k) {
    return (size_t)k * d.sk + (size_t)j * d.sj + (size_t)i;
}

// ------------------------------------------------------------------------
// enthalpy ghost values: top (laser - 2*radiation - evaporation), east/west
// and north/south convection, BoundCondManager.cpp:565-651
// ------------------------------------------------------------------------
struct BcConst {
    double sigm, emiss, htci, htcj, tempamb, pressamb, gasconst, evaplatent, evaptemp, evapfact,
        molarmass;
    int isevap;
};

// The ghost cells the enthalpy boundary kernel sets, numbered g = 0 .. n-1:
// the top face (nkm1) over the interior columns, then the west / east pair
// (i = 0, nim1) of each interior (j, k), then the south / north pair (j = 0,
// njm1) of each interior (i, k).
__device__ __forceinline__ int bc_ghost_count(const Dims &d) {
    return d.njm2 * d.nim2 + 2 * d.nkm2 * d.njm2 + 2 * d.nkm2 * d.nim2;
}
__device__ __forceinline__ size_t bc_ghost_cell(const Dims &d, int g) {
    const int ntop = d.njm2 * d.nim2, nwe = 2 * d.nkm2 * d.njm2;
    if (g < ntop) return cix(d, 1 + g % d.nim2, 1 + g / d.nim2, d.nkm1);
    g -= ntop;
    if (g < nwe) {
        const int r = g >> 1;
        return cix(d, (g & 1) ? d.nim1 : 0, 1 + r % d.njm2, 1 + r / d.njm2);
    }
    g -= nwe;
    const int r = g >> 1;
    return cix(d, 1 + r % d.nim2, (g & 1) ? d.njm1 : 0, 1 + r / d.nim2);
}

// A boundary launch that runs ahead of the convergence test of the
// iteration before it (captured loops): it keeps the ghost values it
// overwrites in save (numbered as above) and sets *spec when it runs (clears
// it when skipped), so that the iteration's later launch can put them back
// if that test ends the loop.
struct BcSave {
    double *save;
    int *spec;
};

struct BcEnthalpyK {
    static constexpr int family = 1, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BcConst k, const double *__restrict__ dxpwinv,
                                                     const double *__restrict__ dypsinv,
                                                     const double *__restrict__ dzpbinv,
                                                     const double *__restrict__ areaij,
                                                     const double *__restrict__ heatin,
                                                     const double *__restrict__ temp,
                                                     const double *__restrict__ diff, double *h,
                                                     double *partial, const int *skip, BcSave sv) {
    __shared__ double scratch[8];
    if (skip && *skip) {
        if (sv.spec && vb.x == 0 && threadIdx.x == 0) *sv.spec = 0;
        return;
    }
    if (sv.spec && vb.x == 0 && threadIdx.x == 0) *sv.spec = 1;
    const int ntop = d.njm2 * d.nim2, new_ = d.nkm2 * d.njm2, nns = d.nkm2 * d.nim2;
    const int t = vb.x * blockDim.x + threadIdx.x;
    double loss = 0.0;
    if (t < ntop) {
        int j = 1 + t / d.nim2, i = 1 + t % d.nim2;
        size_t top = cix(d, i, j, d.nkm1), below = top - d.sk;
        double T = temp[top];
        double hevap = 0.0;
        if (k.isevap) {
            double psat = k.pressamb * libm_exp(-k.evaplatent * k.molarmass / k.gasconst * (1.0 / T - 1.0 / k.evaptemp));
            double mdot = k.evapfact * psat * sqrt(k.molarmass / (2.0 * M_PI * k.gasconst * T));
            hevap = mdot * k.evaplatent;
        }
        double ta = k.tempamb;
        double hrad = k.emiss * k.sigm * (T * T * T * T - ta * ta * ta * ta);
        double c1 = diff[below] * dzpbinv[d.nkm1];
        if (sv.save) sv.save[t] = h[top];
        h[top] = h[below] + (heatin[j * d.ni + i] - hrad - hrad - hevap) / c1;
        loss = (hrad + hrad + hevap) * areaij[j * d.ni + i];
    } else if (t < ntop + new_) {
        int r = t - ntop;
        int kk = 1 + r / d.njm2, j = 1 + r % d.njm2;
        size_t w = cix(d, 0, j, kk), e = cix(d, d.nim1, j, kk);
        if (sv.save) {
            sv.save[ntop + 2 * r] = h[w];
            sv.save[ntop + 2 * r + 1] = h[e];
        }
        double q = k.htci * (temp[w] - k.tempamb);
        double c1 = diff[w + 1] * dxpwinv[1];
        h[w] = h[w + 1] - q / c1;
        q = k.htci * (temp[e] - k.tempamb);
        c1 = diff[e - 1] * dxpwinv[d.nim2];
        h[e] = h[e - 1] - q / c1;
    } else if (t < ntop + new_ + nns) {
        int r = t - ntop - new_;
        int kk = 1 + r / d.nim2, i = 1 + r % d.nim2;
        size_t s = cix(d, i, 0, kk), n = cix(d, i, d.njm1, kk);
        if (sv.save) {
            sv.save[ntop + 2 * new_ + 2 * r] = h[s];
            sv.save[ntop + 2 * new_ + 2 * r + 1] = h[n];
        }
        double q = k.htcj * (temp[s] - k.tempamb);
        double c1 = diff[s + d.sj] * dypsinv[1];
        h[s] = h[s + d.sj] - q / c1;
        q = k.htcj * (temp[n] - k.tempamb);
        c1 = diff[n - d.sj] * dypsinv[d.njm2];
        h[n] = h[n - d.sj] - q / c1;
    }
    double v[1] = {loss};
    const int lane = threadIdx.x;
    v[0] = wave_sum(v[0]);
    if ((lane & 63) == 0) scratch[lane >> 6] = v[0];
    __syncthreads();
    if (lane == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += scratch[w];
        partial[vb.x] = s;
    }
}
};
__global__ __launch_bounds__(256) void k_bc_enthalpy(Dims d, BcConst k, const double *__restrict__ dxpwinv,
                                                     const double *__restrict__ dypsinv,
                                                     const double *__restrict__ dzpbinv,
                                                     const double *__restrict__ areaij,
                                                     const double *__restrict__ heatin,
                                                     const double *__restrict__ temp,
                                                     const double *__restrict__ diff, double *h,
                                                     double *partial, const int *skip, BcSave sv) {
    BcEnthalpyK::run(vb_hw(), d, k, dxpwinv, dypsinv, dzpbinv, areaij, heatin, temp, diff, h, partial, skip, sv);
}

// ------------------------------------------------------------------------
// enthalpy coefficients + sources + boundary folding + relaxation, fused
// with the residual and the block-correction column sums.
//   discretizeEquations(5)  CFDSolverManager.cpp:1157-1263
//   calculateSource(5)      CFDSolverManager.cpp:1677-1843
//   calculateResiduals(5)   CFDSolverManager.cpp:2073-2101
//   enhanceConvergenceSpeed CFDSolverManager.cpp:2305-2332 (column sums)
// Each face coefficient is computed by the cell that owns it (gather), with
// the same operand order as the reference's scatter loop.
// ------------------------------------------------------------------------
struct AsmConst {
    double delt, hlatnt, urfh;
    double avolpow, apowseta, heatthick, heatrb, avolfact, beamx, beamy;
    int pi, scheme;
    int vol;   // mixed mode: the deck has a volumetric source (else the source field stays zero)
};

// Tiles of 64 (i) x 4 (j) columns; each thread marches KA planes up k and
// carries the (k-1, k, k+1) values of the fields read at k-neighbours in
// registers, so every input array is read about once.  Workgroups are
// dispatched to the 8 XCDs round-robin by linear id; the tiles are
// renumbered so that each XCD takes one contiguous run of them (x, then y,
// then z), which keeps the j-neighbour rows of a tile in that XCD's L2.
constexpr int KA = 8;
#ifndef MX_ASM_WAVES
#define MX_ASM_WAVES 2
#endif

__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb / 8, r = nb % 8, x = b % 8, idx = b / 8;
    return (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Output arrays of the assembly.  FP64 mode: the nine coefficient arrays and
// the per-cell block-correction terms.  Mixed mode (MX): the seven matrix
// coefficients rounded to FP32, su in FP64, and per-tile column partials of
// the block-correction sums, part[q][tile][i] (q = bl, blp, blm, blc; tile =
// the (j, k) tile index), formed from the FP32-rounded coefficients so that
// the correction matches the system the sweeps solve.
struct AsmOut {
    double *an, *as, *ae, *aw, *at, *ab, *ap, *su, *src, *part_res, *term_bl, *term_blc;
    float *fan, *fas, *fae, *faw, *fat, *fab, *fap;
    double *blkpart;
    double *h_out;   // the enthalpy, where the late split launch restores ghost values
};
// the mixed assembly split around the flow part of the iteration before
// (captured loops): mode 0 every tile, 1 the tiles away from the device box,
// 2 the tiles near it, which also restores the ghost values kept in gsave
// when the loop has ended and the iteration's boundary launch ran (gspec)
struct AsmSplit {
    const int *dbox;
    int mode;
    const double *gsave;
    const int *gspec;
    int tgx, tgy, tgz;   // mode 1 with a smaller grid: the tile grid its workgroups stride over
};

template <int SCH>
__device__ __forceinline__ void energy_assemble_body(
    Dims d, AsmConst q, DevGeom g, const double *__restrict__ uvel, const double *__restrict__ vvel,
    const double *__restrict__ wvel, const double *__restrict__ rho, const double *__restrict__ diff,
    const double *__restrict__ hnot, const double *__restrict__ fracl, const double *__restrict__ fraclnot,
    const double *__restrict__ h, AsmOut o) {
    __shared__ double scratch[2 * 4];
    double *an_o = o.an, *as_o = o.as, *ae_o = o.ae, *aw_o = o.aw, *at_o = o.at, *ab_o = o.ab, *ap_o = o.ap;
    double *su_o = o.su, *src_o = o.src, *part_res = o.part_res, *term_bl = o.term_bl, *term_blc = o.term_blc;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int gx = gridDim.x, gy = gridDim.y;
    const int nb = gx * gy * gridDim.z;
    const int t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), nb);
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int i = 1 + bx * TX + tx;
    const int j = 1 + by * TY + ty;
    const int k0 = 1 + bz * KA;
    const bool col_ok = (i <= d.nim2) && (j <= d.njm2) && (k0 <= d.nkm2);
    double sumd = 0.0, sumh = 0.0;
    // the scheme is a template constant: the other schemes' weight functions
    // (the exponential's exp) cost no registers
    constexpr int sch = SCH;
    const double ztop = g.z[d.nkm1];
    if (col_ok) {
        const double fxi = g.fracx[i], fxw = g.fracx[i - 1], fyj = g.fracy[j], fys = g.fracy[j - 1];
        const double dxc = g.xu[i + 1] - g.xu[i], dyc = g.yv[j + 1] - g.yv[j];
        const double aij = g.areaij[j * d.ni + i];
        const double xb = q.beamx - g.x[i], yb = q.beamy - g.y[j];
        const int kend = min(k0 + KA - 1, d.nkm2);
        // values at k-neighbours, carried from plane to plane
        size_t c = cix(d, i, j, k0);
        double dB = diff[c - d.sk], dP = diff[c];
        double rP = rho[c], wb = wvel[c];
        double flB = fracl[c - d.sk], flP = fracl[c];
        double hB = h[c - d.sk], hP = h[c];
        for (int k = k0; k <= kend; ++k, c += d.sk) {
            const double dT = diff[c + d.sk], rT = rho[c + d.sk], wt = wvel[c + d.sk];
            const double flT = fracl[c + d.sk], hT = h[c + d.sk];
            // same-plane neighbours, all inside the ghost layer, read up front
            const double dN = diff[c + d.sj], dS = diff[c - d.sj], dE = diff[c + 1], dW = diff[c - 1];
            const double rN = rho[c + d.sj], rE = rho[c + 1];
            const double vs = vvel[c], vn = vvel[c + d.sj], uw = uvel[c], ue = uvel[c + 1];
            const double flE = fracl[c + 1], flW = fracl[c - 1], flN = fracl[c + d.sj], flS = fracl[c - d.sj];
            const double hN = h[c + d.sj], hS = h[c - d.sj], hE = h[c + 1], hW = h[c - 1];
            const double hnP = hnot[c], flnP = fraclnot[c];
            const double aik = g.areaik[k * d.ni + i], ajk = g.areajk[k * d.nj + j];
            const double vol = dxc * dyc * (g.zw[k + 1] - g.zw[k]);
            double an, as, ae, aw, at, ab;
            // north face
            if (j < d.njm2) {
                double fn = rN * vn * aik;
                double dif = dP * dN / ((1.0 - fyj) * dN + fyj * dP);
                an = coef_minus(sch, dif * aik * g.dypsinv[j + 1], fn);
            } else {
                double fn = rN * vn * aik;
                an = coef_minus(sch, dP * aik * g.dypsinv[d.njm1], fn);
            }
            // south face
            {
                double fs = rP * vs * aik;
                if (j > 1) {
                    double dif = dS * dP / ((1.0 - fys) * dP + fys * dS);
                    as = coef_plus(sch, dif * aik * g.dypsinv[j], fs);
                } else {
                    as = coef_plus(sch, dP * aik * g.dypsinv[1], fs);
                }
            }
            // east face
            if (i < d.nim2) {
                double fe = rE * ue * ajk;
                double dif = dP * dE / ((1.0 - fxi) * dE + fxi * dP);
                ae = coef_minus(sch, dif * ajk * g.dxpwinv[i + 1], fe);
            } else {
                double fe = rE * ue * ajk;
                ae = coef_minus(sch, dP * ajk * g.dxpwinv[d.nim1], fe);
            }
            // west face
            {
                double fw = rP * uw * ajk;
                if (i > 1) {
                    double dif = dW * dP / ((1.0 - fxw) * dP + fxw * dW);
                    aw = coef_plus(sch, dif * ajk * g.dxpwinv[i], fw);
                } else {
                    aw = coef_plus(sch, dP * ajk * g.dxpwinv[1], fw);
                }
            }
            // top face
            if (k < d.nkm2) {
                double ft = rT * wt * aij;
                double fz = g.fracz[k];
                double dif = dP * dT / ((1.0 - fz) * dT + fz * dP);
                at = coef_minus(sch, dif * aij * g.dzpbinv[k + 1], ft);
            } else {
                double ft = rT * wt * aij;
                at = coef_minus(sch, dP * aij * g.dzpbinv[d.nkm1], ft);
            }
            // bottom face
            {
                double fb = rP * wb * aij;
                if (k > 1) {
                    double fz = g.fracz[k - 1];
                    double dif = dB * dP / ((1.0 - fz) * dP + fz * dB);
                    ab = coef_plus(sch, dif * aij * g.dzpbinv[k], fb);
                } else {
                    ab = coef_plus(sch, dP * aij * g.dzpbinv[1], fb);
                }
            }
            const double apnot = rP / q.delt * vol;
            double sp = 0.0;
            double su = apnot * hnP;
            // volumetric heat source
            double src = 0.0;
            if (ztop - g.z[k] <= q.heatthick) {
                src = q.avolpow * q.avolfact / M_PI / (q.heatrb * q.heatrb) / q.heatthick * q.apowseta *
                      libm_exp(-q.avolfact / (q.heatrb * q.heatrb) * ((xb * xb) + (yb * yb))) * (double)q.pi;
                su += vol * src;
            }
            // latent heat: transient and advection terms
            double v1 = q.hlatnt * rP / q.delt;
            double volht = vol * v1;
            su -= volht * (flP - flnP);
            {
                double flew = ajk * (dmax(uw, 0.0) * flW - dmax(-uw, 0.0) * flP +
                                     dmax(-ue, 0.0) * flE - dmax(ue, 0.0) * flP);
                double flns = aik * (dmax(vs, 0.0) * flS - dmax(-vs, 0.0) * flP +
                                     dmax(-vn, 0.0) * flN - dmax(vn, 0.0) * flP);
                double fltb = aij * (dmax(wb, 0.0) * flB - dmax(-wb, 0.0) * flP +
                                     dmax(-wt, 0.0) * flT - dmax(wt, 0.0) * flP);
                su += rP * q.hlatnt * (flew + flns + fltb);
            }
            // fixed boundary values folded into the source (bottom, top, south, north, west, east)
            if (k == 1) { su += ab * hB; sp -= ab; ab = 0.0; }
            if (k == d.nkm2) { su += at * hT; sp -= at; at = 0.0; }
            if (j == 1) { su += as * hS; sp -= as; as = 0.0; }
            if (j == d.njm2) { su += an * hN; sp -= an; an = 0.0; }
            if (i == 1) { su += aw * hW; sp -= aw; aw = 0.0; }
            if (i == d.nim2) { su += ae * hE; sp -= ae; ae = 0.0; }
            double ap = an + as + ae + aw + at + ab + apnot - sp;
            ap /= q.urfh;
            su += (1.0 - q.urfh) * ap * hP;
            an_o[c] = an; as_o[c] = as; ae_o[c] = ae; aw_o[c] = aw; at_o[c] = at; ab_o[c] = ab;
            ap_o[c] = ap; su_o[c] = su; src_o[c] = src;
            // residual (calculateResiduals(5))
            double r = (an * hN + as * hS + ae * hE + aw * hW + at * hT + ab * hB + su) / ap - hP;
            sumd += fabs(r);
            sumh += fabs(hP);
            // per-cell terms of the block-correction column sums (summed per
            // x-column in the reference's (k, j) order by k_column_sums)
            term_bl[c] = ap - an - as - at - ab;
            term_blc[c] = ae * hE + aw * hW + an * hN + as * hS + at * hT + ab * hB + su - ap * hP;
            // the next plane's bottom and centre values
            dB = dP; dP = dT;
            rP = rT; wb = wt;
            flB = flP; flP = flT;
            hB = hP; hP = hT;
        }
    }
    double v[2] = {sumd, sumh};
    block_sum<2>(v, scratch);   // contains __syncthreads
    if (tx == 0 && ty == 0) {
        part_res[2 * t] = v[0];
        part_res[2 * t + 1] = v[1];
    }
}

template <int SCH>
__global__ __launch_bounds__(TX *TY) void k_energy_assemble(
    Dims d, AsmConst q, DevGeom g, const double *__restrict__ uvel, const double *__restrict__ vvel,
    const double *__restrict__ wvel, const double *__restrict__ rho, const double *__restrict__ diff,
    const double *__restrict__ hnot, const double *__restrict__ fracl, const double *__restrict__ fraclnot,
    const double *__restrict__ h, AsmOut o) {
    energy_assemble_body<SCH>(d, q, g, uvel, vvel, wvel, rho, diff, hnot, fracl, fraclnot, h, o);
}

// Mixed mode: the same discretisation with the face coefficients formed in
// FP32 (the harmonic-mean diffusivities, the face fluxes, the weighting
// A(|P|), the neighbour coefficients, ap with its relaxation) and stored in
// FP32; the source su (transient, latent heat, boundary folds, relaxation
// term), the residual and the block-correction partials stay FP64 and use
// the FP32 coefficients as stored, so that they describe the very system the
// line passes solve.  FP32 halves the registers of the coefficient work,
// which raises the waves per SIMD that hide the loads' latency; its divides
// take the hardware reciprocal (1 ulp), as the coefficients are FP32 values
// anyway.  Tiles of 64 (i) x 4 (j) columns, as the FP64 form's: narrower
// tiles idle fewer lanes on x-lines that are not multiples of 64 but write
// more partial cache lines, and measure slower.
constexpr int MTX = 64, MTY = 4;
static_assert(MTY >= 4, "the block-correction fold takes one quantity per j-row of the first four");
__device__ __forceinline__ float fmax_first(float a, float b) { return (a < b) ? b : a; }

template <int SCH>
__device__ __forceinline__ float weight_f32(float pe) {
    if constexpr (SCH == AMCFD_DISC_POWERLAW) {
        const float b = 1.0f - 0.1f * pe;
        const float b2 = b * b;
        return fmax_first(0.0f, b2 * b2 * b);
    } else if constexpr (SCH == AMCFD_DISC_EXPONENTIAL) {
        return pe / (expf(pe) - 1.0f);
    } else {
        return 1.0f;
    }
}
template <int SCH>
__device__ __forceinline__ float coef_plus_f32(float dif, float f) {
    return dif * weight_f32<SCH>(fabsf(f) * __builtin_amdgcn_rcpf(dif)) + fmax_first(0.0f, f);
}
template <int SCH>
__device__ __forceinline__ float coef_minus_f32(float dif, float f) {
    return dif * weight_f32<SCH>(fabsf(f) * __builtin_amdgcn_rcpf(dif)) + fmax_first(0.0f, -f);
}
__device__ __forceinline__ float hmean_f32(float p, float q, float wq, float wp) {
    return p * q * __builtin_amdgcn_rcpf(wq * q + wp * p);
}

template <int SCH>
struct AssembleMxK {
    static constexpr int family = 30 + SCH, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, AsmConst q, DevGeom g, const double *__restrict__ uvel, const double *__restrict__ vvel,
                     const double *__restrict__ wvel, const double *__restrict__ rho,
                     const double *__restrict__ diff, const double *__restrict__ hnot,
                     const double *__restrict__ fracl, const double *__restrict__ fraclnot,
                     const double *__restrict__ h, AsmOut o, const int *__restrict__ skip,
                     const StepDev *__restrict__ step, AsmSplit sp2) {
    if (skip && *skip) {
        // the iteration does not run: the ghost values its boundary launch
        // set ahead of the convergence test go back (the late launch only)
        if (sp2.mode == 2 && sp2.gspec && *sp2.gspec) {
            const int ng = bc_ghost_count(d);
            const int nt = MTX * MTY, tid = threadIdx.x + MTX * threadIdx.y;
            const int nbk = vb.gx * vb.gy * vb.gz, lin0 = vb.x + vb.gx * (vb.y + vb.gy * vb.z);
            for (int gi = lin0 * nt + tid; gi < ng; gi += nbk * nt) o.h_out[bc_ghost_cell(d, gi)] = sp2.gsave[gi];
        }
        return;
    }
    if (step) {   // a captured loop: the step's source parameters from device memory
        const SourceParams sp = step->sp;
        q.avolpow = sp.avolpow; q.apowseta = sp.apowseta; q.heatthick = sp.heatthick; q.heatrb = sp.heatrb;
        q.avolfact = sp.avolfact; q.beamx = sp.beamposx; q.beamy = sp.beamposy; q.pi = sp.powerindicator;
    }
    __shared__ double scratch[2 * 4];
    __shared__ double s_blk[4 * MTY * MTX];
    // plane k of the tile and its one-cell ring (row 0 / MTY + 1: south /
    // north, column 0 / MTX + 1: west / east), double-buffered in k: the
    // fields read at horizontal neighbours, FP32 where the coefficients use
    // them in FP32, the enthalpy in FP64
    __shared__ float s_df[2][MTY + 2][MTX + 2], s_rh[2][MTY + 2][MTX + 2], s_u[2][MTY + 2][MTX + 2];
    __shared__ float s_v[2][MTY + 2][MTX + 2], s_fl[2][MTY + 2][MTX + 2];
    __shared__ double s_h[2][MTY + 2][MTX + 2];
    double cbl = 0.0, cblp = 0.0, cblm = 0.0, cblc = 0.0;   // this column's block-correction partials
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int gx = vb.gx, gy = vb.gy;
    const int nb = gx * gy * vb.gz;
    // a grouped launch has renumbered its blocks for the XCDs already
    const int lin = vb.x + gx * (vb.y + gy * vb.z);
    const int t = vb.hw ? xcd_tile(lin, nb) : lin;
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int i = 1 + bx * MTX + tx;
    const int j = 1 + by * MTY + ty;
    const int k0 = 1 + bz * KA;
    const int kend = min(k0 + KA - 1, d.nkm2);
    if (sp2.mode) {
        // The flow part of the iteration before changes velocities only in
        // [istat, iend] x [jstat, jend + 1] x [kstat, nkm1] of the device box;
        // a tile reads velocities one cell beyond its own cells.  Tiles
        // within two cells of that region (near) wait for the flow part (the
        // late launch, mode 2); the others run alongside it (mode 1).
        const int *b = sp2.dbox;
        const int i0 = 1 + bx * MTX, i1 = i0 + MTX - 1, j0 = 1 + by * MTY, j1 = j0 + MTY - 1;
        const bool near = i1 + 2 >= b[0] && i0 - 2 <= b[1] && j1 + 2 >= b[2] && j0 - 2 <= b[3] + 1 && kend + 2 >= b[4];
        if (near != (sp2.mode == 2)) return;
    }
    const bool col_ok = (i <= d.nim2) && (j <= d.njm2) && (k0 <= d.nkm2);
    // the staged cells: every thread its own (interior or the east / north
    // ghost), and 136 threads one cell of the ring each
    const bool own_ok = i <= d.nim1 && j <= d.njm1;
    int hi = -1, hj = -1, hii = 0, hjj = 0;   // the ring cell of this thread
    if (ty == 0) { hi = i; hj = j - 1; hii = tx + 1; hjj = 0; }
    else if (ty == 1) { hi = i; hj = 1 + by * MTY + MTY; hii = tx + 1; hjj = MTY + 1; }
    else if (ty == 2 && tx < MTY) { hi = 1 + bx * MTX - 1; hj = 1 + by * MTY + tx; hii = 0; hjj = tx + 1; }
    else if (ty == 3 && tx < MTY) { hi = 1 + bx * MTX + MTX; hj = 1 + by * MTY + tx; hii = MTX + 1; hjj = tx + 1; }
    const bool ring_ok = hi >= 0 && hi <= d.nim1 && hj >= 0 && hj <= d.njm1;
    const size_t own_c = own_ok ? cix(d, i, j, 0) : 0, ring_c = ring_ok ? cix(d, hi, hj, 0) : 0;
    // one staged cell's values (FP64 as loaded)
    struct Staged {
        double df, rh, u, v, fl, h;
    };
    auto load_cell = [&](size_t c, Staged &v) {
        v.df = diff[c]; v.rh = rho[c]; v.u = uvel[c]; v.v = vvel[c]; v.fl = fracl[c]; v.h = h[c];
    };
    auto store_cell = [&](int buf, int jj, int ii, const Staged &v) {
        s_df[buf][jj][ii] = (float)v.df; s_rh[buf][jj][ii] = (float)v.rh; s_u[buf][jj][ii] = (float)v.u;
        s_v[buf][jj][ii] = (float)v.v; s_fl[buf][jj][ii] = (float)v.fl; s_h[buf][jj][ii] = v.h;
    };
    double sumd = 0.0, sumh = 0.0;
    // column constants
    float fxi = 0.f, fxw = 0.f, fyj = 0.f, fys = 0.f, ryn = 0.f, rys = 0.f, rxe = 0.f, rxw = 0.f, faij = 0.f;
    double dxy = 0.0, src_col = 0.0, ztop = 0.0;
    if (col_ok) {
        ztop = g.z[d.nkm1];
        fxi = (float)g.fracx[i]; fxw = (float)g.fracx[i - 1];
        fyj = (float)g.fracy[j]; fys = (float)g.fracy[j - 1];
        ryn = (float)g.dypsinv[j < d.njm2 ? j + 1 : d.njm1]; rys = (float)g.dypsinv[j > 1 ? j : 1];
        rxe = (float)g.dxpwinv[i < d.nim2 ? i + 1 : d.nim1]; rxw = (float)g.dxpwinv[i > 1 ? i : 1];
        dxy = (g.xu[i + 1] - g.xu[i]) * (g.yv[j + 1] - g.yv[j]);
        faij = (float)g.areaij[j * d.ni + i];
        // volumetric source: its horizontal Gaussian is the column's own
        if (q.heatthick > 0.0) {
            const double xb = q.beamx - g.x[i], yb = q.beamy - g.y[j];
            src_col = q.avolpow * q.avolfact / M_PI / (q.heatrb * q.heatrb) / q.heatthick * q.apowseta *
                      libm_exp(-q.avolfact / (q.heatrb * q.heatrb) * ((xb * xb) + (yb * yb))) * (double)q.pi;
        }
    }
    const double rdelt = 1.0 / q.delt;
    const float rurfh = 1.0f / (float)q.urfh;
    // the march: plane k from the staged buffer k & 1, its bottom and centre
    // values carried in registers, its top values (plane k + 1) loaded one
    // plane ahead and staged for the next plane
    Staged own{}, ring{};
    double wnext = 0.0, hn_next = 0.0, fln_next = 0.0;
    float dB = 0.f, dP = 0.f, wb = 0.f, flB = 0.f;
    double rP = 0.0, flP = 0.0, hB = 0.0, hP = 0.0;
    if (k0 <= d.nkm2) {
        if (own_ok) {
            const size_t c = own_c + (size_t)k0 * d.sk;
            Staged p0;
            load_cell(c, p0);
            store_cell(k0 & 1, ty + 1, tx + 1, p0);
            dB = (float)diff[c - d.sk]; flB = (float)fracl[c - d.sk]; hB = h[c - d.sk];
            dP = (float)p0.df; rP = p0.rh; wb = (float)wvel[c]; flP = p0.fl; hP = p0.h;
            load_cell(c + d.sk, own);
            wnext = wvel[c + d.sk];
            hn_next = hnot[c];
            fln_next = fraclnot[c];
        }
        if (ring_ok) {
            const size_t c = ring_c + (size_t)k0 * d.sk;
            Staged r0;
            load_cell(c, r0);
            store_cell(k0 & 1, hjj, hii, r0);
            load_cell(c + d.sk, ring);
        }
    }
    __syncthreads();
    for (int k = k0; k <= kend; ++k) {
        // plane k + 1 to its buffer (values loaded last plane); its top
        // values are this plane's top neighbours
        const int cur = k & 1, nxt = cur ^ 1;
        if (own_ok) store_cell(nxt, ty + 1, tx + 1, own);
        if (ring_ok) store_cell(nxt, hjj, hii, ring);
        const float dT = (float)own.df, wt = (float)wnext;
        const double rT = own.rh, flT = own.fl, hT = own.h;
        const double hnP = hn_next, flnP = fln_next;
        // plane k + 2 and this plane's successor's old values, in flight
        // while plane k is computed
        if (k + 1 <= kend) {
            if (own_ok) {
                const size_t c = own_c + (size_t)(k + 1) * d.sk;
                load_cell(c + d.sk, own);
                wnext = wvel[c + d.sk];
                hn_next = hnot[c];
                fln_next = fraclnot[c];
            }
            if (ring_ok) load_cell(ring_c + (size_t)(k + 2) * d.sk, ring);
        }
        if (col_ok) {
            const size_t c = own_c + (size_t)k * d.sk;
            const int jj = ty + 1, ii = tx + 1;
            const float dN = s_df[cur][jj + 1][ii], dS = s_df[cur][jj - 1][ii];
            const float dE = s_df[cur][jj][ii + 1], dW = s_df[cur][jj][ii - 1];
            const float rN = s_rh[cur][jj + 1][ii], rE = s_rh[cur][jj][ii + 1], rPf = (float)rP, rTf = (float)rT;
            const float vs = s_v[cur][jj][ii], vn = s_v[cur][jj + 1][ii], uw = s_u[cur][jj][ii], ue = s_u[cur][jj][ii + 1];
            const float flE = s_fl[cur][jj][ii + 1], flW = s_fl[cur][jj][ii - 1];
            const float flN = s_fl[cur][jj + 1][ii], flS = s_fl[cur][jj - 1][ii];
            const double hN = s_h[cur][jj + 1][ii], hS = s_h[cur][jj - 1][ii], hE = s_h[cur][jj][ii + 1],
                         hW = s_h[cur][jj][ii - 1];
            const float aik = (float)g.areaik[k * d.ni + i], ajk = (float)g.areajk[k * d.nj + j];
            const double vol = dxy * (g.zw[k + 1] - g.zw[k]);
            float an, as, ae, aw, at, ab;
            an = coef_minus_f32<SCH>((j < d.njm2 ? hmean_f32(dP, dN, 1.0f - fyj, fyj) : dP) * aik * ryn, rN * vn * aik);
            as = coef_plus_f32<SCH>((j > 1 ? hmean_f32(dS, dP, 1.0f - fys, fys) : dP) * aik * rys, rPf * vs * aik);
            ae = coef_minus_f32<SCH>((i < d.nim2 ? hmean_f32(dP, dE, 1.0f - fxi, fxi) : dP) * ajk * rxe, rE * ue * ajk);
            aw = coef_plus_f32<SCH>((i > 1 ? hmean_f32(dW, dP, 1.0f - fxw, fxw) : dP) * ajk * rxw, rPf * uw * ajk);
            {
                const float fzt = (float)g.fracz[k], rzt = (float)g.dzpbinv[k < d.nkm2 ? k + 1 : d.nkm1];
                at = coef_minus_f32<SCH>((k < d.nkm2 ? hmean_f32(dP, dT, 1.0f - fzt, fzt) : dP) * faij * rzt,
                                         rTf * wt * faij);
                const float fzb = (float)g.fracz[k - 1], rzb = (float)g.dzpbinv[k > 1 ? k : 1];
                ab = coef_plus_f32<SCH>((k > 1 ? hmean_f32(dB, dP, 1.0f - fzb, fzb) : dP) * faij * rzb,
                                        rPf * wb * faij);
            }
            const double apnot = rP * rdelt * vol;
            float sp = 0.0f;
            double su = apnot * hnP;
            double src = 0.0;
            if (ztop - g.z[k] <= q.heatthick) {
                src = src_col;
                su += vol * src;
            }
            // latent heat: transient (FP64) and advection (FP32) terms
            su -= vol * (q.hlatnt * rP * rdelt) * (flP - flnP);
            {
                const float flPf = (float)flP;
                const float flew = ajk * (fmax_first(uw, 0.0f) * flW - fmax_first(-uw, 0.0f) * flPf +
                                          fmax_first(-ue, 0.0f) * flE - fmax_first(ue, 0.0f) * flPf);
                const float flns = aik * (fmax_first(vs, 0.0f) * flS - fmax_first(-vs, 0.0f) * flPf +
                                          fmax_first(-vn, 0.0f) * flN - fmax_first(vn, 0.0f) * flPf);
                const float fltb = faij * (fmax_first(wb, 0.0f) * flB - fmax_first(-wb, 0.0f) * flPf +
                                           fmax_first(-wt, 0.0f) * (float)flT - fmax_first(wt, 0.0f) * flPf);
                su += rP * q.hlatnt * (double)(flew + flns + fltb);
            }
            // fixed boundary values folded into the source
            if (k == 1) { su += (double)ab * hB; sp -= ab; ab = 0.0f; }
            if (k == d.nkm2) { su += (double)at * hT; sp -= at; at = 0.0f; }
            if (j == 1) { su += (double)as * hS; sp -= as; as = 0.0f; }
            if (j == d.njm2) { su += (double)an * hN; sp -= an; an = 0.0f; }
            if (i == 1) { su += (double)aw * hW; sp -= aw; aw = 0.0f; }
            if (i == d.nim2) { su += (double)ae * hE; sp -= ae; ae = 0.0f; }
            const float ap = (an + as + ae + aw + at + ab + (float)apnot - sp) * rurfh;
            su += (1.0 - q.urfh) * (double)ap * hP;
            o.fan[c] = an; o.fas[c] = as; o.fae[c] = ae; o.faw[c] = aw; o.fat[c] = at; o.fab[c] = ab;
            o.fap[c] = ap;
            o.su[c] = su;
            if (q.vol) o.src[c] = src;
            const double rap = ap;
            const double num = (double)ae * hE + (double)aw * hW + (double)an * hN + (double)as * hS +
                               (double)at * hT + (double)ab * hB + su;
            sumd += fabs(num / rap - hP);
            sumh += fabs(hP);
            cbl += rap - (double)an - (double)as - (double)at - (double)ab;
            cblp += (double)ae;
            cblm += (double)aw;
            cblc += num - rap * hP;
        }
        // the next plane's bottom and centre values
        dB = dP; dP = dT;
        rP = rT; wb = wt;
        flB = (float)flP; flP = flT;
        hB = hP; hP = hT;
        __syncthreads();
    }
    // column partials of this tile: the j-rows of each i, in order
    s_blk[(0 * MTY + ty) * MTX + tx] = cbl;
    s_blk[(1 * MTY + ty) * MTX + tx] = cblp;
    s_blk[(2 * MTY + ty) * MTX + tx] = cblm;
    s_blk[(3 * MTY + ty) * MTX + tx] = cblc;
    __syncthreads();
    if (ty < 4 && i <= d.nim2) {
        const int ntile = gy * vb.gz, tile = by + gy * bz;
        const int qq = ty;   // one quantity per j-row of the first four
        double a = 0.0;
        for (int r = 0; r < MTY; ++r) a += s_blk[(qq * MTY + r) * MTX + tx];
        o.blkpart[((size_t)qq * ntile + tile) * d.ni + i] = a;
    }
    double v[2] = {sumd, sumh};
    block_sum<2>(v, scratch);   // contains __syncthreads
    if (tx == 0 && ty == 0) {
        o.part_res[2 * t] = v[0];
        o.part_res[2 * t + 1] = v[1];
    }
}
};
template <int SCH>
__global__ __launch_bounds__(MTX *MTY) __attribute__((amdgpu_waves_per_eu(MX_ASM_WAVES, 8))) void k_energy_assemble_mx(Dims d, AsmConst q, DevGeom g, const double *__restrict__ uvel, const double *__restrict__ vvel,
                     const double *__restrict__ wvel, const double *__restrict__ rho,
                     const double *__restrict__ diff, const double *__restrict__ hnot,
                     const double *__restrict__ fracl, const double *__restrict__ fraclnot,
                     const double *__restrict__ h, AsmOut o, const int *__restrict__ skip,
                     const StepDev *__restrict__ step, AsmSplit sp2) {
    AssembleMxK<SCH>::run(vb_hw(), d, q, g, uvel, vvel, wvel, rho, diff, hnot, fracl, fraclnot, h, o, skip, step, sp2);
}
// fewer workgroups than tiles (room for the launches beside it): each
// strides over the tile grid sp2.tgx x tgy x tgz
template <int SCH>
__global__ __launch_bounds__(MTX *MTY) __attribute__((amdgpu_waves_per_eu(MX_ASM_WAVES, 8))) void k_energy_assemble_mx_strided(Dims d, AsmConst q, DevGeom g, const double *__restrict__ uvel, const double *__restrict__ vvel,
                     const double *__restrict__ wvel, const double *__restrict__ rho,
                     const double *__restrict__ diff, const double *__restrict__ hnot,
                     const double *__restrict__ fracl, const double *__restrict__ fraclnot,
                     const double *__restrict__ h, AsmOut o, const int *__restrict__ skip,
                     const StepDev *__restrict__ step, AsmSplit sp2) {
    const int nt = sp2.tgx * sp2.tgy * sp2.tgz;
    for (int v = blockIdx.x; v < nt; v += gridDim.x) {
        const VB vb{v % sp2.tgx, (v / sp2.tgx) % sp2.tgy, v / (sp2.tgx * sp2.tgy), sp2.tgx, sp2.tgy, sp2.tgz, 1};
        AssembleMxK<SCH>::run(vb, d, q, g, uvel, vvel, wvel, rho, diff, hnot, fracl, fraclnot, h, o, skip, step, sp2);
        __syncthreads();   // the tile's LDS is reused by the next
    }
}

// ------------------------------------------------------------------------
// phase-wise diffusivity / density / viscosity, CFDSolverManager.cpp:2137-2265
// ------------------------------------------------------------------------
struct PropConst {
    double tliquid, tsolid, viscos, dens, denl, acpa, acpb, acpl, ka, kb, kc, kla, klb;
    double pden, pcpa, pcpb, pka, pkb;
    int powder;
};

// the properties of node t from its temperature, liquid fraction and
// powder fraction (CFDSolverManager.cpp:624 and the property routines it calls)
__device__ __forceinline__ void props_at(const PropConst &c, size_t t, double T, const double *__restrict__ fracl,
                                         const double *__restrict__ csfrac, double &dv, double &rv, double &vv) {
    double diffl = (c.kla * T + c.klb) / c.acpl;
    dv = diffl;
    rv = c.denl;
    vv = c.viscos;
    if (!(T >= c.tliquid)) {
        double diffs, rhos;
        if (!c.powder) {
            diffs = (c.ka * (T * T) + c.kb * T + c.kc) / (c.acpa * T + c.acpb);
            rhos = c.dens;
        } else {
            double cs = csfrac[t];
            double cps = c.acpa * T + c.acpb;
            double conds = c.ka * (T * T) + c.kb * T + c.kc;
            double diffp = (c.pka * T + c.pkb) / (c.pcpa * T + c.pcpb);
            diffs = (1.0 - cs) * diffp + cs * conds / cps;
            rhos = (1.0 - cs) * c.pden + cs * c.dens;
        }
        dv = diffs;
        rv = rhos;
        vv = 1e10;
        if (!(T <= c.tsolid)) {
            double fl = fracl[t];
            dv = fl * diffl + (1.0 - fl) * diffs;
            rv = fl * c.denl + (1.0 - fl) * rhos;
            vv = c.viscos;
        }
    }
}

struct PropsK {
    static constexpr int family = 3, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, size_t n, PropConst c, const double *__restrict__ temp,
                                               const double *__restrict__ fracl,
                                               const double *__restrict__ csfrac, double *diff, double *rho,
                                               double *vis, const int *skip) {
    if (skip && *skip) return;
    size_t t = (size_t)vb.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double dv, rv, vv;
    props_at(c, t, temp[t], fracl, csfrac, dv, rv, vv);
    diff[t] = dv;
    rho[t] = rv;
    vis[t] = vv;
}
};
__global__ __launch_bounds__(256) void k_props(size_t n, PropConst c, const double *__restrict__ temp,
                                               const double *__restrict__ fracl,
                                               const double *__restrict__ csfrac, double *diff, double *rho,
                                               double *vis, const int *skip) {
    PropsK::run(vb_hw(), n, c, temp, fracl, csfrac, diff, rho, vis, skip);
}

// ------------------------------------------------------------------------
// block correction: fold column sums, solve the 1-D system along x and
// (second kernel) add the correction, CFDSolverManager.cpp:2270-2362.
// Also folds the energy residual, CFDSolverManager.cpp:2099.
// ------------------------------------------------------------------------
// Column sums bl, blp, blm, blc for every x index i, each accumulated in the
// reference's order (k outer, j inner).  Floating-point sums are order
// dependent and the SIMPLE iteration is chaotic enough that a reordered sum
// changes the trajectory after a few dozen steps, so the order is kept.  Each
// sum is one dependent chain of (nj-2)(nk-2) adds; to keep the chains fed,
// one workgroup serves CS_COLS columns of one quantity: lanes 0..CS_COLS-1 of
// wave 0 hold the running sums while the other waves stream the next CS_CH
// rows into an LDS double buffer.  Output: blk[q*ni + i], q = bl, blp, blm,
// blc.  CS_NT = 1024 threads with 128 KB of LDS is the form for a solver
// alone; beside other solvers a 256-thread form with 32 KB finds a CU sooner
// (the sums are the same chains in the same order).
constexpr int CS_COLS = 16;

template <int CS_NT, int CS_CH>
__global__ __launch_bounds__(CS_NT) void k_column_sums(Dims d, const double *__restrict__ term_bl,
                                                       const double *__restrict__ ae,
                                                       const double *__restrict__ aw,
                                                       const double *__restrict__ term_blc, double *blk) {
    __shared__ double buf[2][CS_CH * CS_COLS];
    constexpr int NL = CS_NT - 64;                          // loader threads
    constexpr int PER = (CS_CH * CS_COLS + NL - 1) / NL;    // elements per loader per chunk
    const int q = blockIdx.y;
    const double *src = q == 0 ? term_bl : q == 1 ? ae : q == 2 ? aw : term_blc;
    const int i0 = 1 + blockIdx.x * CS_COLS;
    const int R = d.njm2 * d.nkm2;
    const int nch = (R + CS_CH - 1) / CS_CH;
    const int tid = threadIdx.x;
    const bool loader = tid >= 64;
    auto load_chunk = [&](int c, double *dst) {
        const int lt = tid - 64;
        double v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = lt + u * NL;
            const int r = c * CS_CH + e / CS_COLS, i = i0 + e % CS_COLS;
            v[u] = 0.0;
            if (e < CS_CH * CS_COLS && r < R && i <= d.nim2) {
                const int k = 1 + r / d.njm2, j = 1 + r % d.njm2;
                v[u] = src[(size_t)k * d.sk + (size_t)j * d.sj + i];
            }
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = lt + u * NL;
            if (e < CS_CH * CS_COLS) dst[e] = v[u];
        }
    };
    if (loader) load_chunk(0, buf[0]);
    __syncthreads();
    double s = 0.0;
    for (int c = 0; c < nch; ++c) {
        if (loader) {
            if (c + 1 < nch) load_chunk(c + 1, buf[(c + 1) & 1]);
        } else if (tid < CS_COLS) {
            const double *b = buf[c & 1] + tid;
            const int rows = min(CS_CH, R - c * CS_CH);
            int r = 0;
            for (; r + 16 <= rows; r += 16) {
                double t[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) t[u] = b[(r + u) * CS_COLS];
#pragma unroll
                for (int u = 0; u < 16; ++u) s += t[u];
            }
            for (; r < rows; ++r) s += b[r * CS_COLS];
        }
        __syncthreads();
    }
    if (tid < CS_COLS && i0 + tid <= d.nim2) blk[(size_t)q * d.ni + i0 + tid] = s;
}

// 1-D correction along x by the reference's recurrence (one lane, exact
// order), CFDSolverManager.cpp:2334-2349.  The column sums are staged in LDS
// first so the recurrence never waits on global memory.
constexpr int BS_MAX = 1032;   // ni of the longest x-line handled, rounded up

__global__ __launch_bounds__(256) void k_block_solve(Dims d, const double *__restrict__ blk, double *delphi) {
    __shared__ double bl[BS_MAX], blp[BS_MAX], blm[BS_MAX], blc[BS_MAX], pib[BS_MAX], qib[BS_MAX];
    for (int i = threadIdx.x; i < d.ni; i += 256) {
        bl[i] = blk[i];
        blp[i] = blk[d.ni + i];
        blm[i] = blk[2 * d.ni + i];
        blc[i] = blk[3 * d.ni + i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double pv = blp[1] / bl[1], qv = blc[1] / bl[1];
        pib[1] = pv;
        qib[1] = qv;
        for (int i = 2; i < d.nim1; ++i) {   // the running values stay in registers
            const double den = bl[i] - blm[i] * pv;
            pv = blp[i] / den;
            qv = (blc[i] + blm[i] * qv) / den;
            pib[i] = pv;
            qib[i] = qv;
        }
        double nxt = qv;
        qib[d.nim2] = nxt;
        for (int i = d.nim2 - 1; i > 0; --i) {
            nxt = pib[i] * nxt + qib[i];
            qib[i] = nxt;
        }
    }
    __syncthreads();
    for (int i = 1 + threadIdx.x; i <= d.nim2; i += 256) delphi[i] = qib[i];
}

__global__ __launch_bounds__(256) void k_add_delphi(Dims d, const double *__restrict__ delphi, double *phi) {
    const int i = 1 + blockIdx.x * 64 + threadIdx.x;
    const int j = 1 + blockIdx.y * 4 + threadIdx.y;
    if (i > d.nim2 || j > d.njm2) return;
    const double dp = delphi[i];
    for (int k = 1 + blockIdx.z * KB; k < 1 + (blockIdx.z + 1) * KB && k <= d.nkm2; ++k)
        phi[cix(d, i, j, k)] += dp;
}

// ------------------------------------------------------------------------
// enthalpy -> temperature / liquid fraction on every active node
// (CFDSolverManager.cpp:2531-2571), fused with the peak temperature and the
// heat-balance terms of BoundCondManager.cpp:746-844.
// ------------------------------------------------------------------------
struct H2tConst {
    double hlcal, hsmelt, acpa, acpb, acpl, tliquid, tsolid, hlatnt, delt;
    int steady;
};

// PROPS: the node's properties are computed first, from the temperature and
// liquid fraction the node holds on entry, stored, and used by the heat-balance
// terms in place of the stored ones (the same values, one read less); no node
// reads another node's properties, so the fusion keeps every result.
template <bool PROPS>
__device__ __forceinline__ void h2t_body(const VB &vb, const Dims &d, const H2tConst &q, const PropConst &pc,
                                         const DevGeom &g, const double *__restrict__ h,
                                         const double *__restrict__ hnot, const double *__restrict__ fraclnot,
                                         const double *__restrict__ csfrac, const double *__restrict__ src,
                                         double *__restrict__ temp, double *__restrict__ fracl,
                                         double *__restrict__ diff, double *__restrict__ rho,
                                         double *__restrict__ vis, double *partial, const int *skip) {
    __shared__ double scratch[10 * 4];
    if (skip && *skip) return;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i = vb.x * TX + tx;
    const int j = vb.y * TY + ty;
    double tmax = -INFINITY;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // accul heatvol W E B T S N
    const double dT = q.tliquid - q.tsolid;
    if (i < d.ni && j < d.nj) {
        // every plane's inputs are loaded before any is used, so a thread has
        // its KB planes' loads in flight at once
        double hpre[KB], tpre[KB], hnpre[KB], fnpre[KB], rpre[KB];
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
            const int k = vb.z * KB + kk;
            if (k >= d.nk) break;
            const size_t c = cix(d, i, j, k);
            hpre[kk] = h[c];
            if constexpr (PROPS) tpre[kk] = temp[c];
            if (!q.steady) {
                hnpre[kk] = hnot[c];
                fnpre[kk] = fraclnot[c];
                if constexpr (!PROPS) rpre[kk] = rho[c];
            }
        }
#pragma unroll
        for (int kk = 0; kk < KB; ++kk) {
            const int k = vb.z * KB + kk;
            if (k >= d.nk) break;
            const size_t c = cix(d, i, j, k);
            double dv = 0.0, rv = 0.0;
            if constexpr (PROPS) {
                double vv;
                props_at(pc, c, tpre[kk], fracl, csfrac, dv, rv, vv);
                diff[c] = dv;
                rho[c] = rv;
                vis[c] = vv;
            } else {
                rv = rpre[kk];
            }
            const double hv = hpre[kk];
            double T, fl;
            if (hv >= q.hlcal) {
                fl = 1.0;
                T = (hv - q.hlcal) / q.acpl + q.tliquid;
            } else if (hv <= q.hsmelt) {
                fl = 0.0;
                T = (sqrt(q.acpb * q.acpb + 2 * q.acpa * hv) - q.acpb) / q.acpa;
            } else {
                fl = (hv - q.hsmelt) / (q.hlcal - q.hsmelt);
                T = dT * fl + q.tsolid;
            }
            temp[c] = T;
            fracl[c] = fl;
            tmax = dmax(tmax, T);
            const bool in = i >= 1 && i <= d.nim2 && j >= 1 && j <= d.njm2 && k >= 1 && k <= d.nkm2;
            if (!in) continue;
            if (!q.steady) {
                double vol = (g.xu[i + 1] - g.xu[i]) * (g.yv[j + 1] - g.yv[j]) * (g.zw[k + 1] - g.zw[k]);
                double dh = hv - hnpre[kk] + (fl - fnpre[kk]) * q.hlatnt;
                acc[0] += vol * rv * dh / q.delt;
                if (src) acc[1] += src[c] * vol;   // null: no volumetric source, the field is zero
            }
            if (i == 1) acc[2] += g.areajk[k * d.nj + j] * ((PROPS ? dv : diff[c]) * (h[c - 1] - hv) * g.dxpwinv[1]);
            if (i == d.nim2) acc[3] += g.areajk[k * d.nj + j] * ((PROPS ? dv : diff[c]) * (h[c + 1] - hv) * g.dxpwinv[d.nim1]);
            if (k == 1) acc[4] += g.areaij[j * d.ni + i] * ((PROPS ? dv : diff[c]) * (h[c - d.sk] - hv) * g.dzpbinv[1]);
            if (k == d.nkm2) acc[5] += g.areaij[j * d.ni + i] * ((PROPS ? dv : diff[c]) * (h[c + d.sk] - hv) * g.dzpbinv[d.nkm1]);
            if (j == 1) acc[6] += g.areaik[k * d.ni + i] * ((PROPS ? dv : diff[c]) * (h[c - d.sj] - hv) * g.dypsinv[1]);
            if (j == d.njm2) acc[7] += g.areaik[k * d.ni + i] * ((PROPS ? dv : diff[c]) * (h[c + d.sj] - hv) * g.dypsinv[d.njm1]);
        }
    }
    // block reduction: max of T, sums of the rest
    const int lane = tx + ty * TX, nw = (TX * TY) / 64;
    tmax = wave_max(tmax);
#pragma unroll
    for (int qq = 0; qq < 8; ++qq) acc[qq] = wave_sum(acc[qq]);
    if ((lane & 63) == 0) {
        scratch[lane >> 6] = tmax;
#pragma unroll
        for (int qq = 0; qq < 8; ++qq) scratch[(qq + 1) * nw + (lane >> 6)] = acc[qq];
    }
    __syncthreads();
    if (lane == 0) {
        const int b = vb.x + vb.gx * (vb.y + vb.gy * vb.z);
        double m = scratch[0];
        for (int w = 1; w < nw; ++w) m = dmax(m, scratch[w]);
        partial[(size_t)b * 10] = m;
        for (int qq = 0; qq < 8; ++qq) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += scratch[(qq + 1) * nw + w];
            partial[(size_t)b * 10 + 1 + qq] = s;
        }
    }
}

struct H2tK {
    static constexpr int family = 4, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, H2tConst q, DevGeom g, const double *__restrict__ h,
                                                 const double *__restrict__ hnot,
                                                 const double *__restrict__ fraclnot,
                                                 const double *__restrict__ rho,
                                                 const double *__restrict__ diff,
                                                 const double *__restrict__ src, double *temp, double *fracl,
                                                 double *partial, const int *skip) {
        h2t_body<false>(vb, d, q, PropConst{}, g, h, hnot, fraclnot, nullptr, src, temp, fracl,
                        const_cast<double *>(diff), const_cast<double *>(rho), nullptr, partial, skip);
    }
};

// the properties (CFDSolverManager.cpp:624) and H -> T in one pass, for loops
// where nothing between the two reads or writes what they read or write
struct PropsH2tK {
    static constexpr int family = 40, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, H2tConst q, PropConst pc, DevGeom g,
                                               const double *__restrict__ h, const double *__restrict__ hnot,
                                               const double *__restrict__ fraclnot,
                                               const double *__restrict__ csfrac, const double *__restrict__ src,
                                               double *temp, double *fracl, double *diff, double *rho, double *vis,
                                               double *partial, const int *skip) {
        h2t_body<true>(vb, d, q, pc, g, h, hnot, fraclnot, csfrac, src, temp, fracl, diff, rho, vis, partial, skip);
    }
};
__global__ __launch_bounds__(TX *TY) void k_h2t(Dims d, H2tConst q, DevGeom g, const double *__restrict__ h,
                                                 const double *__restrict__ hnot,
                                                 const double *__restrict__ fraclnot,
                                                 const double *__restrict__ rho,
                                                 const double *__restrict__ diff,
                                                 const double *__restrict__ src, double *temp, double *fracl,
                                                 double *partial, const int *skip) {
    H2tK::run(vb_hw(), d, q, g, h, hnot, fraclnot, rho, diff, src, temp, fracl, partial, skip);
}

__global__ __launch_bounds__(TX *TY) void k_props_h2t(Dims d, H2tConst q, PropConst pc, DevGeom g,
                                                       const double *__restrict__ h, const double *__restrict__ hnot,
                                                       const double *__restrict__ fraclnot,
                                                       const double *__restrict__ csfrac,
                                                       const double *__restrict__ src, double *temp, double *fracl,
                                                       double *diff, double *rho, double *vis, double *partial,
                                                       const int *skip) {
    PropsH2tK::run(vb_hw(), d, q, pc, g, h, hnot, fraclnot, csfrac, src, temp, fracl, diff, rho, vis, partial, skip);
}

// ------------------------------------------------------------------------
// peak temperature, pool length / depth / width and the momentum box,
// BoundCondManager.cpp:280-447 (one workgroup; every scan of the reference
// is a first-crossing or extremum search, done here as parallel min/max).
// ------------------------------------------------------------------------
constexpr int MP_MAXNI = 4100;   // x-nodes of the longest grid a line solver takes
struct MeltPoolK {
    static constexpr int family = 5, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, int nblk, const double *__restrict__ partial,
                                                    const double *__restrict__ temp,
                                                    const double *__restrict__ x, const double *__restrict__ y,
                                                    const double *__restrict__ z, double ts, int istart,
                                                    int jstart, DevScalars *sc, int ctl, int navierstokes,
                                                    const StepDev *step) {
    __shared__ int si[5];
    if (ctl && sc->done) return;
    if (step) {   // the beam's nearest node of this step (captured loops)
        istart = step->istart;
        jstart = step->jstart;
    }
    __shared__ double sd[32];
    __shared__ double tp_s;
    __shared__ int s_kf[MP_MAXNI], s_ju[MP_MAXNI], s_jd[MP_MAXNI], s_rd[MP_MAXNI], s_rw[MP_MAXNI];
    __shared__ int s_nd, s_nw;
    const int tid = threadIdx.x, nt = blockDim.x;
    // peak temperature (first-argument-wins max over block partials)
    double m = -INFINITY;
    {   // each thread's blocks in order, eight blocks' loads in flight at a time
        constexpr int U = 8;
        int b = tid;
        for (; b + (U - 1) * nt < nblk; b += U * nt) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = partial[(size_t)(b + u * nt) * 10];
#pragma unroll
            for (int u = 0; u < U; ++u) m = dmax(m, v[u]);
        }
        for (; b < nblk; b += nt) m = dmax(m, partial[(size_t)b * 10]);
    }
    m = wave_max(m);
    if ((tid & 63) == 0) sd[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
        double mm = sd[0];
        for (int w = 1; w < nt / 64; ++w) mm = dmax(mm, sd[w]);
        tp_s = mm;
        sc->tpeak = mm;
        if (ctl) sc->noflow = (mm > ts && navierstokes) ? 0 : 1;   // CFDSolverManager.cpp:660
    }
    __syncthreads();
    const double tpeak = tp_s;
    if (!(tpeak > ts)) {
        if (tid == 0) { sc->alen = 0.0; sc->depth = 0.0; sc->width = 0.0; }
        return;
    }
    // The width scan can step one row past the last interior row (as the
    // reference does); such reads see the neighbouring memory of the field
    // when it exists and 0 otherwise, never outside the allocation.
    const long long total = (long long)d.nkf * (long long)d.sk;
    auto T = [&](int i, int j, int k) {
        long long idx = (long long)k * (long long)d.sk + (long long)j * (long long)d.sj + i;
        return (idx >= 0 && idx < total) ? temp[idx] : 0.0;
    };
    auto Y = [&](int j) { return (j >= 0 && j < d.nj) ? y[j] : 0.0; };
    const int js = jstart, top = d.nkm1;
    // One pass over row js: the length (first cell at or below ts from the
    // beam on, last one below ts before it) and the rows that start a depth
    // run (at nkm2) or a width run (on the top plane).  A row whose starting
    // cell is below ts has no run (it leaves the fold unchanged); for the
    // others, the run ends at the first cell below ts:
    //   kf[i] = highest k < nkm2 with T(i, js, k) < ts (0 if none),
    //   ju[i] = lowest j in (js, njm2] with T(i, j, top) < ts (njm1 if none),
    //   jd[i] = highest j in [1, js) with T(i, j, top) < ts (0 if none),
    // and the run covers k in [kf + 1, nkm2], j in [js, ju - 1], j in
    // [jd + 1, js]; the folds below are the loops' min / max.  Only the rows
    // with a run are scanned, all their cells at once, eight loads in flight
    // per thread.
    if (tid == 0) {
        s_nd = 0;
        s_nw = 0;
        si[0] = 0x7fffffff;
        si[1] = -0x7fffffff;
    }
    __syncthreads();
    {   // the length scans span istart..nim2 and 1..istart (istart is 0..nim1)
        int cmax = d.nim2, cmin = 1;
        for (int i = tid; i <= d.nim1; i += nt) {
            const double tt = T(i, js, top);
            if (i >= istart && i <= d.nim2 && tt <= ts) cmax = min(cmax, i);
            if (i >= 1 && i <= istart && tt < ts) cmin = max(cmin, i);
            if (i < 1 || i > d.nim2) continue;
            const bool dr = !(T(i, js, d.nkm2) < ts), wr = !(tt < ts);
            s_kf[i] = dr ? 0 : d.nkm2;
            s_ju[i] = wr ? d.njm1 : js;
            s_jd[i] = wr ? 0 : js;
            if (dr) s_rd[atomicAdd(&s_nd, 1)] = i;
            if (wr) s_rw[atomicAdd(&s_nw, 1)] = i;
        }
        atomicMin(&si[0], cmax);
        atomicMax(&si[1], cmin);
    }
    __syncthreads();
    const int imax = si[0], imin = si[1];
    const int nrd = s_nd, nrw = s_nw;
    auto scan = [&](int total, auto cell, auto hit) {
        constexpr int U = 8;
        for (int e0 = tid; e0 < total; e0 += U * nt) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * nt;
                v[u] = e < total ? cell(e) : ts;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int e = e0 + u * nt;
                if (e < total && v[u] < ts) hit(e);
            }
        }
    };
    const int nkd = d.nkm2 - 1, nju = d.njm2 - js, njd = js - 1;
    scan(nrd * nkd, [&](int e) { return T(s_rd[e % nrd], js, 1 + e / nrd); },
         [&](int e) { atomicMax(&s_kf[s_rd[e % nrd]], 1 + e / nrd); });
    if (nju > 0)
        scan(nrw * nju, [&](int e) { return T(s_rw[e % nrw], js + 1 + e / nrw, top); },
             [&](int e) { atomicMin(&s_ju[s_rw[e % nrw]], js + 1 + e / nrw); });
    if (njd > 0)
        scan(nrw * njd, [&](int e) { return T(s_rw[e % nrw], 1 + e / nrw, top); },
             [&](int e) { atomicMax(&s_jd[s_rw[e % nrw]], 1 + e / nrw); });
    __syncthreads();
    int ck = d.nkm2, cju = js, cjd = js;
    for (int i = 1 + tid; i <= d.nim2; i += nt) {
        ck = min(ck, s_kf[i] + 1);
        cju = max(cju, s_ju[i] - 1);
        cjd = min(cjd, s_jd[i] + 1);
    }
    if (tid == 0) {   // slots of their own: si[0..1] are still read above
        si[2] = 0x7fffffff;
        si[3] = -0x7fffffff;
        si[4] = 0x7fffffff;
    }
    __syncthreads();
    atomicMin(&si[2], ck);
    atomicMax(&si[3], cju);
    atomicMin(&si[4], cjd);
    __syncthreads();
    const int kmin = si[2] - 1, jmax = si[3] + 1, jmin = si[4] - 1;
    // interpolated extents (max / min over i of per-row values)
    // the three interpolations of each row in one pass (each folds as its
    // own loop would)
    double dep = -INFINITY, ymx = -INFINITY, ymn = INFINITY;
    const bool kd = kmin != 0, ju = jmax != js, jd = jmin != js;
    for (int i = 1 + tid; i <= d.nim2; i += nt) {
        if (kd && !(T(i, js, kmin + 1) < ts)) {
            double dtdz = (z[kmin] - z[kmin - 1]) / (T(i, js, kmin) - T(i, js, kmin - 1));
            dep = dmax(z[top] - z[kmin] + (T(i, js, kmin) - ts) * dtdz, dep);
        }
        if (ju && !(T(i, jmax - 1, top) < ts)) {
            double dtdy = (Y(jmax) - Y(jmax + 1)) / (T(i, jmax, top) - T(i, jmax + 1, top));
            ymx = dmax(Y(jmax) + (ts - T(i, jmax, top)) * dtdy, ymx);
        }
        if (jd && !(T(i, jmin + 1, top) < ts)) {
            double dtdy = (Y(jmin) - Y(jmin - 1)) / (T(i, jmin, top) - T(i, jmin - 1, top));
            ymn = dmin(Y(jmin) + (ts - T(i, jmin, top)) * dtdy, ymn);
        }
    }
    dep = wave_max(dep);
    ymx = wave_max(ymx);
    ymn = -wave_max(-ymn);
    __syncthreads();
    if ((tid & 63) == 0) {
        sd[tid >> 6] = dep;
        sd[16 + (tid >> 6)] = ymx;
    }
    __syncthreads();
    double depth_all = -INFINITY, ymax_all = -INFINITY;
    if (tid == 0) {
        for (int w = 0; w < nt / 64; ++w) {
            depth_all = dmax(sd[w], depth_all);
            ymax_all = dmax(sd[16 + w], ymax_all);
        }
    }
    __syncthreads();
    if ((tid & 63) == 0) sd[tid >> 6] = ymn;
    __syncthreads();
    if (tid == 0) {
        double ymin_all = INFINITY;
        for (int w = 0; w < nt / 64; ++w) ymin_all = dmin(sd[w], ymin_all);
        // length from the interpolated crossing on each side
        double dx = (x[imax] - x[imax + 1]) / (T(imax, js, top) - T(imax + 1, js, top));
        double xxmax = x[imax] + (ts - T(imax, js, top)) * dx;
        dx = (x[imin] - x[imin - 1]) / (T(imin, js, top) - T(imin - 1, js, top));
        double xxmin = x[imin] + (ts - T(imin, js, top)) * dx;
        sc->alen = xxmax - xxmin;
        double depth = 0.0;
        if (kmin == 0) depth = z[top] - z[0];
        else depth = dmax(depth_all, 0.0);
        sc->depth = depth;
        double yymax = y[0];
        if (jmax != js) yymax = dmax(ymax_all, yymax);
        double yymin = y[js];
        if (jmin != js) yymin = dmin(ymin_all, yymin);
        sc->width = yymax - yymin;
        sc->imin = imin;
        sc->imax = imax;
        sc->jmin = jmin;
        sc->jmax = jmax;
        sc->kmin = kmin;
        sc->istat = max(imin - 3, 1);
        sc->iend = min(imax + 3, d.nim2);
        sc->jstat = max(jmin - 3, 1);
        sc->jend = min(jmax + 2, d.njm2);
        sc->kstat = max(kmin - 2, 2);
    }
}
};
__global__ __launch_bounds__(1024) void k_melt_pool(Dims d, int nblk, const double *__restrict__ partial,
                                                    const double *__restrict__ temp,
                                                    const double *__restrict__ x, const double *__restrict__ y,
                                                    const double *__restrict__ z, double ts, int istart,
                                                    int jstart, DevScalars *sc, int ctl, int navierstokes,
                                                    const StepDev *step) {
    MeltPoolK::run(vb_hw(), d, nblk, partial, temp, x, y, z, ts, istart, jstart, sc, ctl, navierstokes, step);
}

// ------------------------------------------------------------------------
// end-of-iteration scalars: residual folds, heat balance, energy ratio
// (CFDSolverManager.cpp:1989-2071, BoundCondManager.cpp:846-857)
// ------------------------------------------------------------------------
// the convergence test of CFDSolverManager.cpp:690-716 on the device, after
// the folds (captured loops): counts the iteration; the first largest of
// resorm, resoru, resorv, resorw; the heat-balance band with the laser on
// (or in steady mode), resorh < 1e-6 with it off; done at maxit
__device__ __forceinline__ void converge_test(DevScalars *sc, const StepDev *step, int steady, double tol) {
    sc->niter += 1;
    double res[4] = {sc->resorm, sc->resoru, sc->resorv, sc->resorw};
    double amaxres = res[0];
    for (int r = 1; r < 4; ++r)
        if (amaxres < res[r]) amaxres = res[r];
    const double ratio = sc->ratio;
    bool ok;
    if (steady) ok = ratio <= 1.001 && ratio >= 0.999 && amaxres < tol;
    else if (step->powerindicator == 1) ok = ratio <= 1.01 && ratio >= 0.99 && amaxres < tol;
    else ok = sc->resorh < 1e-6 && amaxres < tol;
    if (ok || sc->niter >= step->maxit) {
        sc->converged = ok ? 1 : 0;
        sc->done = 1;
        sc->noflow = 1;
    }
}

// a[q] += p[b * stride + off + q] for q < NQV over b = tid, tid + nt, ... < n,
// in that order (the sums' bits are those of the plain loop), with the loads
// of U blocks in flight at once: the partials were just written and sit in
// the caches, so the plain loop is a chain of cache latencies
template <int NQV, int U>
__device__ __forceinline__ void strided_fold(const double *__restrict__ p, int n, int stride, int off, int tid,
                                             int nt, double *a) {
    int b = tid;
    for (; b + (U - 1) * nt < n; b += U * nt) {
        double v[U][NQV];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NQV; ++q) v[u][q] = p[(size_t)(b + u * nt) * stride + off + q];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NQV; ++q) a[q] += v[u][q];
    }
    for (; b < n; b += nt)
#pragma unroll
        for (int q = 0; q < NQV; ++q) a[q] += p[(size_t)b * stride + off + q];
}

struct FinalizeK {
    static constexpr int family = 6, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, int nblk_t, int nblk_b, int nblk_f, int nblk_e, int nblk_c,
                                                  const double *__restrict__ part_t,
                                                  const double *__restrict__ part_b,
                                                  const double *__restrict__ part_f,
                                                  const double *__restrict__ part_e,
                                                  const double *__restrict__ part_c, int flow_done, int species_done,
                                                  int steady, double dens, double width, DevScalars *sc, int ctl,
                                                  const StepDev *step, double tol) {
    if (ctl) {   // the loop's state lives on the device
        if (sc->done) return;
        flow_done = !sc->noflow;
        width = sc->width;
    }
    // every quantity here only feeds the convergence test, so the partials
    // are folded by a parallel tree (fixed order: still reproducible)
    constexpr int NQ = 17;   // 8 heat-balance, toploss, 2 energy residual, 4 flow, 2 species (+ umax)
    __shared__ double s_red[NQ + 1][4];
    const int tid = threadIdx.x, nt = blockDim.x;
    double a[NQ + 1];
#pragma unroll
    for (int q = 0; q <= NQ; ++q) a[q] = 0.0;
    strided_fold<8, 8>(part_t, nblk_t, 10, 1, tid, nt, a);
    strided_fold<1, 8>(part_b, nblk_b, 1, 0, tid, nt, a + 8);
    strided_fold<2, 8>(part_e, nblk_e, 2, 0, tid, nt, a + 9);
    if (species_done) strided_fold<2, 8>(part_c, nblk_c, 2, 0, tid, nt, a + 15);
    if (flow_done) {
        for (int b = tid; b < nblk_f; b += nt) {
            const double *pu = part_f + (size_t)b * 3, *pv = part_f + ((size_t)nblk_f + b) * 3;
            const double *pw = part_f + ((size_t)2 * nblk_f + b) * 3, *pp = part_f + ((size_t)3 * nblk_f + b) * 3;
            a[11] += pu[0];
            a[NQ] = dmax(a[NQ], pu[1]);
            a[12] += pv[0];
            a[13] += pw[0];
            a[14] += pp[1];           // continuity normaliser
        }
    }
    double mass = 0.0;   // sum of |mass imbalance| (pressure-correction partials)
    if (flow_done)
        for (int b = tid; b < nblk_f; b += nt) mass += part_f[((size_t)3 * nblk_f + b) * 3 + 2];
#pragma unroll
    for (int q = 0; q < NQ; ++q) a[q] = wave_sum(a[q]);
    a[NQ] = wave_max(a[NQ]);
    mass = wave_sum(mass);
    __shared__ double s_mass[4];
    if ((tid & 63) == 0) {
#pragma unroll
        for (int q = 0; q <= NQ; ++q) s_red[q][tid >> 6] = a[q];
        s_mass[tid >> 6] = mass;
    }
    __syncthreads();
    if (tid != 0) return;
    double r[NQ + 1];
    for (int q = 0; q < NQ; ++q) r[q] = s_red[q][0] + s_red[q][1] + s_red[q][2] + s_red[q][3];
    r[NQ] = dmax(dmax(s_red[NQ][0], s_red[NQ][1]), dmax(s_red[NQ][2], s_red[NQ][3]));
    const double msum = s_mass[0] + s_mass[1] + s_mass[2] + s_mass[3];
    sc->resorh = r[9] / (r[10] + AMCFD_SMALL);
    sc->accul = steady ? 0.0 : r[0];
    sc->heatvol = steady ? 0.0 : r[1];
    sc->fluxwest = r[2];
    sc->fluxeast = r[3];
    sc->fluxbottom = r[4];
    sc->fluxtop = r[5];
    sc->fluxsouth = r[6];
    sc->fluxnorth = r[7];
    sc->ahtoploss = r[8];
    double heatout = (sc->fluxnorth + sc->fluxbottom + sc->fluxwest + sc->fluxeast + sc->fluxsouth) - sc->ahtoploss;
    sc->heatout = heatout;
    if (steady) sc->ratio = -heatout / sc->heatinlaser;
    else sc->ratio = (sc->heatvol + sc->heatinlaser) / (sc->accul - heatout);
    if (flow_done) {
        double refmom = 0.25 * M_PI * (width * width) * dens * (r[NQ] * r[NQ]);
        sc->refmom = refmom;
        sc->resoru = r[11] / refmom;
        sc->resorv = r[12] / refmom;
        sc->resorw = r[13] / refmom;
        sc->resorm = msum / (r[14] * dens + AMCFD_SMALL);
    }
    if (species_done) sc->resorc = r[15] / (r[16] + AMCFD_SMALL);
    if (ctl && step) converge_test(sc, step, steady, tol);   // the same thread: it reads what it wrote
}
};
__global__ __launch_bounds__(256) void k_finalize(int nblk_t, int nblk_b, int nblk_f, int nblk_e, int nblk_c,
                                                  const double *__restrict__ part_t,
                                                  const double *__restrict__ part_b,
                                                  const double *__restrict__ part_f,
                                                  const double *__restrict__ part_e,
                                                  const double *__restrict__ part_c, int flow_done, int species_done,
                                                  int steady, double dens, double width, DevScalars *sc, int ctl,
                                                  const StepDev *step, double tol) {
    FinalizeK::run(vb_hw(), nblk_t, nblk_b, nblk_f, nblk_e, nblk_c, part_t, part_b, part_f, part_e, part_c, flow_done,
                   species_done, steady, dens, width, sc, ctl, step, tol);
}

// ------------------------------------------------------------------------
// end-of-step push-back, CFDSolverManager.cpp:2625-2687
// ------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_end_step(size_t n, double tsolid, double tliquid, int powder, double track,
                                                  double trackid, const double *__restrict__ temp,
                                                  const double *__restrict__ h, const double *__restrict__ fracl,
                                                  double *u, double *v, double *w, double *unot, double *vnot,
                                                  double *wnot, double *hnot, double *fraclnot, double *csfrac,
                                                  double *solfrac, double *mpphase, double *mpliq) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double T = temp[t];
    double uu = u[t], vv = v[t], ww = w[t];
    if (T <= tsolid) {
        uu = 0.0; vv = 0.0; ww = 0.0;
        u[t] = 0.0; v[t] = 0.0; w[t] = 0.0;
    }
    unot[t] = uu;
    vnot[t] = vv;
    wnot[t] = ww;
    hnot[t] = h[t];
    double fl = fracl[t];
    fraclnot[t] = fl;
    if (powder) {
        csfrac[t] = dmax(csfrac[t], fl);
        solfrac[t] = dmax(solfrac[t], fl * track);
        if (T > tsolid) mpphase[t] = trackid;
        if (T > tliquid) mpliq[t] = trackid;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_bc_enthalpy(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                        const amcfd_params &p, const int *skip, double *save, int *spec) {
    BcConst k{p.sigm, p.emiss, p.htci, p.htcj, p.tempamb, p.pressamb, p.gasconst, p.evaplatent,
              p.evaptemp, p.evapfact, p.molarmass, p.isevap};
    int n = d.njm2 * d.nim2 + d.nkm2 * d.njm2 + d.nkm2 * d.nim2;
    int nb = (n + 255) / 256;
    dispatch<BcEnthalpyK, 256>(k_bc_enthalpy, dim3(nb), dim3(256), 0, st, d, k, g.dxpwinv, g.dypsinv, g.dzpbinv,
                       g.areaij, f.heatin, f.temp, f.diff, f.enthalpy, pr.bc, skip, BcSave{save, spec});
}

// Gaussian surface flux map, BoundCondManager.cpp:128-159: one thread per
// (i, j); also writes areaij * heatin for the total below
__global__ __launch_bounds__(256) void k_laser_map(int ni, int n, const double *__restrict__ x,
                                                   const double *__restrict__ y,
                                                   const double *__restrict__ areaij, LaserParams q,
                                                   double *__restrict__ heatin, double *__restrict__ contrib) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const int j = t / ni, i = t - j * ni;
    const double ydist = q.beamy - y[j];
    const double xdist = q.beamx - x[i];
    const double dist2 = (xdist * xdist) + (ydist * ydist);
    double h = q.peakhin * libm_exp(-q.alasfact * dist2 / q.rb2) * q.pind;
    if (q.alasfact < 1.0e-6 && dist2 < q.rb2) h = q.varlas / (M_PI * q.rb2) * q.pind;
    heatin[t] = h;
    contrib[t] = areaij[t] * h;
}

// heatinlaser = sum of areaij * heatin in the reference's (j, i) order
// (BoundCondManager.cpp:158): one dependent add per cell.  The chain is kept
// fed from LDS: wave 1 streams chunks of the values into a double buffer
// (coalesced loads, many in flight) while lane 0 of wave 0 adds the chunk
// before it, so the chain waits on LDS reads issued ahead of it instead of
// on memory round trips.
constexpr int LT_CHUNK = 1024;
__global__ __launch_bounds__(128) void k_laser_total(int n, const double *__restrict__ contrib,
                                                     DevScalars *sc) {
    __shared__ double buf[2][LT_CHUNK];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nch = (n + LT_CHUNK - 1) / LT_CHUNK;
    auto stage = [&](int ch) {
        double *b = buf[ch & 1];
#pragma unroll
        for (int q = 0; q < LT_CHUNK / 64; ++q) {
            const int r = q * 64 + lane, c = ch * LT_CHUNK + r;
            b[r] = c < n ? contrib[c] : 0.0;
        }
    };
    if (wave == 1 && nch > 0) stage(0);
    __syncthreads();
    double total = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
        if (wave == 1) {
            if (ch + 1 < nch) stage(ch + 1);
        } else if (lane == 0) {
            const double *b = buf[ch & 1];
            const int m = min(LT_CHUNK, n - ch * LT_CHUNK);
            int r = 0;
            for (; r + 32 <= m; r += 32) {
                double v[32];
#pragma unroll
                for (int u = 0; u < 32; ++u) v[u] = b[r + u];
#pragma unroll
                for (int u = 0; u < 32; ++u) total += v[u];
            }
            for (; r < m; ++r) total += b[r];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) sc->heatinlaser = total;
}

void launch_laser_map(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                      const LaserParams &q, DevScalars *sc) {
    const int n = d.ni * d.nj;
    hipLaunchKernelGGL(k_laser_map, dim3((n + 255) / 256), dim3(256), 0, st, d.ni, n, g.x, g.y, g.areaij, q,
                       f.heatin, pr.laser);
    hipLaunchKernelGGL(k_laser_total, dim3(1), dim3(128), 0, st, n, pr.laser, sc);
}

static dim3 energy_grid(const Dims &d) {
    return dim3((d.nim2 + TX - 1) / TX, (d.njm2 + TY - 1) / TY, (d.nkm2 + KB - 1) / KB);
}

static dim3 assemble_grid(const Dims &d) {
    return dim3((d.nim2 + TX - 1) / TX, (d.njm2 + TY - 1) / TY, (d.nkm2 + KA - 1) / KA);
}

static dim3 assemble_grid_mx(const Dims &d) {
    return dim3((d.nim2 + MTX - 1) / MTX, (d.njm2 + MTY - 1) / MTY, (d.nkm2 + KA - 1) / KA);
}

void launch_energy_assemble(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                            const amcfd_params &p, const SourceParams &s) {
    AsmConst q{p.delt, p.hlatnt, p.urfh, s.avolpow, s.apowseta, s.heatthick, s.heatrb, s.avolfact,
               s.beamposx, s.beamposy, s.powerindicator, p.discretization};
    AsmOut o{f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su, f.sourceinput, pr.energy, f.sp, f.apnot,
             nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    auto kern = p.discretization == AMCFD_DISC_POWERLAW     ? k_energy_assemble<AMCFD_DISC_POWERLAW>
                : p.discretization == AMCFD_DISC_EXPONENTIAL ? k_energy_assemble<AMCFD_DISC_EXPONENTIAL>
                                                             : k_energy_assemble<AMCFD_DISC_UPWIND>;
    hipLaunchKernelGGL(kern, assemble_grid(d), dim3(TX, TY), 0, st, d, q, g, f.uvel, f.vvel, f.wvel, f.rho, f.diff,
                       f.hnot, f.fracl, f.fraclnot, f.enthalpy, o);
}

void launch_energy_assemble_mixed(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                                  const MxFields &m, const amcfd_params &p, const SourceParams &s, const int *skip,
                                  const StepDev *step, const int *split_box, int split_mode, const double *gsave,
                                  const int *gspec, int split_blocks) {
    AsmConst q{p.delt, p.hlatnt, p.urfh, s.avolpow, s.apowseta, s.heatthick, s.heatrb, s.avolfact,
               s.beamposx, s.beamposy, s.powerindicator, p.discretization, mixed_has_volumetric_source(p)};
    AsmOut o{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, m.su, f.sourceinput, pr.energy,
             nullptr, nullptr, m.an, m.as, m.ae, m.aw, m.at, m.ab, m.ap, m.blkpart, f.enthalpy};
    const dim3 tg = assemble_grid_mx(d);
    AsmSplit sp2{split_box, split_box ? split_mode : 0, gsave, gspec, 0, 0, 0};
    dim3 grid = tg;
    if (sp2.mode == 1 && split_blocks > 0 && (unsigned)split_blocks < tg.x * tg.y * tg.z) {
        sp2.tgx = (int)tg.x;
        sp2.tgy = (int)tg.y;
        sp2.tgz = (int)tg.z;
        grid = dim3((unsigned)split_blocks);
    }
#define ASM(SCH)                                                                                                   \
    do {                                                                                                           \
        if (sp2.tgx) {                                                                                             \
            if (LaunchRecorder *r = tls_recorder()) r->unsupported = true;                                         \
            else                                                                                                   \
                hipLaunchKernelGGL(k_energy_assemble_mx_strided<SCH>, grid, dim3(MTX, MTY), 0, st, d, q, g, f.uvel,  \
                                   f.vvel, f.wvel, f.rho, f.diff, f.hnot, f.fracl, f.fraclnot, f.enthalpy, o, skip, \
                                   step, sp2);                                                                     \
        } else                                                                                                     \
            dispatch<AssembleMxK<SCH>, MTX * MTY>(k_energy_assemble_mx<SCH>, grid, dim3(MTX, MTY), 0, st, d, q, g, \
                                                   f.uvel, f.vvel, f.wvel, f.rho, f.diff, f.hnot, f.fracl,         \
                                                   f.fraclnot, f.enthalpy, o, skip, step, sp2);                    \
    } while (0)
    if (p.discretization == AMCFD_DISC_POWERLAW) ASM(AMCFD_DISC_POWERLAW);
    else if (p.discretization == AMCFD_DISC_EXPONENTIAL) ASM(AMCFD_DISC_EXPONENTIAL);
    else ASM(AMCFD_DISC_UPWIND);
#undef ASM
}

// A fixed volumetric source of 0 W (finite factors, positive radius and
// thickness) adds exactly zero, as no source does; heat-source tables may
// change the source during the run.
int mixed_has_volumetric_source(const amcfd_params &p) {
    if (p.n_heat_table > 0 || p.volheatsourcetable) return 1;
    if (!p.volheatsource) return 0;
    const bool zero = p.avolpow == 0.0 && p.heatrb > 0.0 && p.heatthick > 0.0 && std::isfinite(p.avolfact) &&
                      std::isfinite(p.apowseta);
    return zero ? 0 : 1;
}

int mixed_tile_count(const Dims &d) {
    const dim3 a = assemble_grid_mx(d);
    return (int)(a.y * a.z);
}

static PropConst prop_const(const amcfd_params &p) {
    return PropConst{p.tliquid, p.tsolid, p.viscos, p.dens, p.denl, p.acpa, p.acpb, p.acpl, p.thconsa,
                     p.thconsb, p.thconsc, p.thconla, p.thconlb, p.powderden, p.powdercpa, p.powdercpb,
                     p.powderthcona, p.powderthconb, p.powderbed};
}

void launch_props(hipStream_t st, const Dims &d, DevFields f, const amcfd_params &p, const int *skip) {
    PropConst c = prop_const(p);
    size_t n = (size_t)d.ni * d.nj * d.nk;
    dispatch<PropsK, 256>(k_props, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, c, f.temp, f.fracl,
                       f.csfrac, f.diff, f.rho, f.vis, skip);
}

void launch_block_correction(hipStream_t st, const Dims &d, DevFields f, Partials pr, double *phi, int shared) {
    const dim3 grid((d.nim2 + CS_COLS - 1) / CS_COLS, 4);
    if (shared)
        hipLaunchKernelGGL((k_column_sums<256, 128>), grid, dim3(256), 0, st, d, f.sp, f.ae, f.aw, f.apnot, pr.blk);
    else
        hipLaunchKernelGGL((k_column_sums<1024, 512>), grid, dim3(1024), 0, st, d, f.sp, f.ae, f.aw, f.apnot, pr.blk);
    hipLaunchKernelGGL(k_block_solve, dim3(1), dim3(256), 0, st, d, pr.blk, pr.delphi);
    hipLaunchKernelGGL(k_add_delphi, energy_grid(d), dim3(64, 4), 0, st, d, pr.delphi, phi);
}

static dim3 node_grid(const Dims &d) {
    return dim3((d.ni + TX - 1) / TX, (d.nj + TY - 1) / TY, (d.nk + KB - 1) / KB);
}

void launch_h2t(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr, const amcfd_params &p,
                const int *skip, bool src_zero) {
    H2tConst q{p.hlcal, p.hsmelt, p.acpa, p.acpb, p.acpl, p.tliquid, p.tsolid, p.hlatnt, p.delt, p.steady};
    dispatch<H2tK, TX * TY>(k_h2t, node_grid(d), dim3(TX, TY), 0, st, d, q, g, f.enthalpy, f.hnot, f.fraclnot, f.rho,
                       f.diff, src_zero ? nullptr : f.sourceinput, f.temp, f.fracl, pr.h2t, skip);
}

void launch_props_h2t(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                      const amcfd_params &p, const int *skip, bool src_zero) {
    H2tConst q{p.hlcal, p.hsmelt, p.acpa, p.acpb, p.acpl, p.tliquid, p.tsolid, p.hlatnt, p.delt, p.steady};
    dispatch<PropsH2tK, TX * TY>(k_props_h2t, node_grid(d), dim3(TX, TY), 0, st, d, q, prop_const(p), g, f.enthalpy,
                                 f.hnot, f.fraclnot, f.csfrac, src_zero ? nullptr : f.sourceinput, f.temp, f.fracl,
                                 f.diff, f.rho, f.vis, pr.h2t, skip);
}

void launch_melt_pool(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                      const amcfd_params &p, int istart, int jstart, DevScalars *sc, bool ctl,
                      const StepDev *step) {
    dim3 gr = node_grid(d);
    int nblk = gr.x * gr.y * gr.z;
    dispatch<MeltPoolK, 1024>(k_melt_pool, dim3(1), dim3(1024), 0, st, d, nblk, pr.h2t, f.temp, g.x, g.y, g.z, p.tsolid,
                       istart, jstart, sc, (int)ctl, p.navierstokes, step);
}

void launch_finalize(hipStream_t st, const Dims &d, Partials pr, DevScalars *sc, const amcfd_params &p,
                     int flow_done, int species_done, int steady, double width, bool ctl, const StepDev *step) {
    dim3 gr = node_grid(d);
    int nblk_t = gr.x * gr.y * gr.z;
    const int nblk_b = (d.njm2 * d.nim2 + d.nkm2 * d.njm2 + d.nkm2 * d.nim2 + 255) / 256;
    const dim3 e = assemble_grid(d), c = energy_grid(d);   // enthalpy and species assembly partials
    dispatch<FinalizeK, 256>(k_finalize, dim3(1), dim3(256), 0, st, nblk_t, nblk_b, pr.nblk_f, (int)(e.x * e.y * e.z),
                       (int)(c.x * c.y * c.z), pr.h2t, pr.bc, pr.flow, pr.energy, pr.species, flow_done, species_done, steady, p.dens,
                       width, sc, (int)ctl, step, p.nonlintol);
}

__global__ void k_loop_reset(DevScalars *sc) {
    sc->niter = 0;
    sc->done = 0;
    sc->converged = 0;
    sc->noflow = 1;
}

// grouped runs: whether every instance's loop is done, into host memory
__global__ __launch_bounds__(64) void k_group_done(int *const *__restrict__ done, int n, int *out) {
    int all = 1;
    for (int i = threadIdx.x; i < n; i += 64) all &= *done[i] != 0;
    all = __all(all);
    if (threadIdx.x == 0) *out = all;
}

void launch_group_done(hipStream_t st, int *const *done, int n, int *out) {
    hipLaunchKernelGGL(k_group_done, dim3(1), dim3(64), 0, st, done, n, out);
}

void launch_loop_reset(hipStream_t st, DevScalars *sc) {
    hipLaunchKernelGGL(k_loop_reset, dim3(1), dim3(1), 0, st, sc);
}

void launch_end_step(hipStream_t st, const Dims &d, DevFields f, const amcfd_params &p, double track,
                     double trackid) {
    size_t n = (size_t)d.ni * d.nj * d.nk;
    hipLaunchKernelGGL(k_end_step, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p.tsolid, p.tliquid,
                       p.powderbed, track, trackid, f.temp, f.enthalpy, f.fracl, f.uvel, f.vvel, f.wvel, f.unot,
                       f.vnot, f.wnot, f.hnot, f.fraclnot, f.csfrac, f.solfrac, f.mpphase, f.mpliqphase);
}

// the device exp (fn 0) or log (fn 1) of n values, for tests against the C
// library: the kernels evaluate exactly these functions
__global__ __launch_bounds__(256) void k_libm_eval(int fn, size_t n, const double *__restrict__ x,
                                                   double *__restrict__ y) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = fn == 0 ? libm_exp(x[t]) : libm_log(x[t]);
}

int eval_libm_device(int fn, const double *x, double *y, size_t n) {
    if (n == 0) return 0;
    double *dx = nullptr, *dy = nullptr;
    hipError_t e = hipMalloc((void **)&dx, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&dy, n * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_libm_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, n, dx, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(y, dy, n * sizeof(double), hipMemcpyDeviceToHost);
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return e == hipSuccess ? 0 : -1;
}

int energy_partial_sizes(const Dims &d, int *nblk_e, int *nrow_e, int *nblk_t) {
    // sized for the larger of the enthalpy (assemble_grid) and species
    // (energy_grid) assembly partials
    dim3 e = energy_grid(d), t = node_grid(d), a = assemble_grid(d);
    *nblk_e = std::max(e.x * e.y * e.z, a.x * a.y * a.z);
    *nrow_e = e.y * e.z;
    *nblk_t = t.x * t.y * t.z;
    return 0;
}

}  // namespace amcfd
