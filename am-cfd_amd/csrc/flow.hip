// Momentum (u, v, w) and pressure-correction kernels on the melt-pool box
// (FP64 plain mode).  Box = k in [kstat, nkm2], j in [jstat, jend],
// i in [istat+1, iend-1] (BoundCondManager.cpp:438-446); blocks are 64 (i) x
// 4 (j) threads, one box plane per grid z-slice.

#include "boxprep.h"
#include "fv.h"

namespace amcfd {

namespace {

constexpr int TX = 64, TY = 4;

struct BoxDev {
    int istat, iend, jstat, jend, kstat;
};

__device__ __forceinline__ size_t cix(const Dims &d, int i, int j, int k) {
    return (size_t)k * d.sk + (size_t)j * d.sj + (size_t)i;
}

// The box a kernel works on: the launch value, or (with dbox) the one the
// melt-pool kernel left in device memory (istat, iend, jstat, jend, kstat),
// for launches whose grid was fixed before the box was known (captured
// graphs).  Its tiles are TX (i) x TY (j) x 1 (k); workgroup b takes tiles
// b, b + G, ...  With G equal to the tile count (the FP64 mode's launches)
// workgroup b takes exactly tile b.
struct BoxTiles {
    BoxDev b;
    int gx, gy, n;
};
__device__ __forceinline__ BoxTiles box_tiles(const Dims &d, BoxDev b, const int *dbox) {
    if (dbox) b = BoxDev{dbox[0], dbox[1], dbox[2], dbox[3], dbox[4]};
    const int ni_b = b.iend - b.istat - 1, nj_b = b.jend - b.jstat + 1, nk_b = d.nkm1 - b.kstat;
    BoxTiles t;
    t.b = b;
    t.gx = (ni_b + TX - 1) / TX;
    t.gy = (nj_b + TY - 1) / TY;
    t.n = (ni_b < 1 || nj_b < 1 || nk_b < 1) ? 0 : t.gx * t.gy * nk_b;
    return t;
}

// zero face velocities next to solid cells, CFDSolverManager.cpp:2577-2620
struct CleanK {
    static constexpr int family = 15, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BoxDev b0, double ts, const double *__restrict__ temp,
                                                   double *u, double *v, double *w, const int *dbox,
                                                   const int *skip) {
    if (skip && *skip) return;
    const BoxTiles bt = box_tiles(d, b0, dbox);
    const BoxDev b = bt.b;
    for (int t = vb.lin(); t < bt.n; t += vb.count()) {
        const int i = b.istat + 1 + (t % bt.gx) * TX + threadIdx.x;
        const int j = b.jstat + ((t / bt.gx) % bt.gy) * TY + threadIdx.y;
        const int k = b.kstat + t / (bt.gx * bt.gy);
        if (i > b.iend - 1 || j > b.jend) continue;
        const size_t c = cix(d, i, j, k);
        const double tP = temp[c];
        if (dmin(tP, temp[c + 1]) <= ts) u[c + 1] = 0.0;
        if (dmin(tP, temp[c + d.sj]) <= ts) v[c + d.sj] = 0.0;
        if (dmin(tP, temp[c + d.sk]) <= ts) w[c + d.sk] = 0.0;
    }
}
};
__global__ __launch_bounds__(TX *TY) void k_clean(Dims d, BoxDev b0, double ts, const double *__restrict__ temp,
                                                   double *u, double *v, double *w, const int *dbox,
                                                   const int *skip) {
    CleanK::run(vb_hw(), d, b0, ts, temp, u, v, w, dbox, skip);
}

// Fe-O surfactant surface-tension gradient, BoundCondManager.cpp:706-741
__device__ double surface_tension_gradient(double t, double a0, double dgdtp) {
    if (a0 < 0) return dgdtp;
    const double A = 4.3e-4, R = 8314, Ts = 2.03e-8, k1 = 1.38e-2, H0 = -1.46e8;
    double K = k1 * libm_exp(-H0 / R / t);
    double g = (-A - R * Ts * libm_log(1 + K * a0) - (K * a0 / (1 + K * a0)) * (Ts * H0 / t)) * 1;
    if (t >= 2500) g = 0;
    return g;
}

// Marangoni top ghost (ivar 1: u, 2: v) and zeroed box sides (ivar 1..4),
// BoundCondManager.cpp:472-564
struct BcFlowK {
    static constexpr int family = 16, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BoxDev b0, int ivar, double a_oxygen, double dgdtp,
                                                 const double *__restrict__ dxpwinv,
                                                 const double *__restrict__ dypsinv,
                                                 const double *__restrict__ dzpbinv,
                                                 const double *__restrict__ fracx,
                                                 const double *__restrict__ fracy,
                                                 const double *__restrict__ temp,
                                                 const double *__restrict__ fracl,
                                                 const double *__restrict__ vis, double *phi, const int *dbox,
                                                 const int *skip) {
    if (skip && *skip) return;
    const BoxDev b = dbox ? BoxDev{dbox[0], dbox[1], dbox[2], dbox[3], dbox[4]} : b0;
    const int ni_b = b.iend - b.istat - 1, nj_b = b.jend - b.jstat + 1, nk_b = d.nkm1 - b.kstat;
    if (ni_b < 1 || nj_b < 1) return;
    const int ntop = (ivar <= 2) ? ni_b * nj_b : 0;
    const int nall = ntop + (nk_b > 0 ? nk_b * nj_b : 0);
    for (int t = vb.x * blockDim.x + threadIdx.x; t < nall; t += vb.gx * blockDim.x) {
    if (t < ntop) {
        const int j = b.jstat + t / ni_b, i = b.istat + 1 + t % ni_b;
        const size_t top = cix(d, i, j, d.nkm1), below = top - d.sk;
        double grad, flf, visf;
        if (ivar == 1) {
            const double fx = fracx[i - 1];
            grad = (temp[top] - temp[top - 1]) * dxpwinv[i];
            flf = fracl[top] * (1.0 - fx) + fracl[top - 1] * fx;
            double vp = vis[below], vw = vis[below - 1];
            visf = vp * vw / (vp * (1.0 - fx) + vw * fx);
        } else {
            const double fy = fracy[j - 1];
            grad = (temp[top] - temp[top - d.sj]) * dypsinv[j];
            flf = fracl[top] * (1.0 - fy) + fracl[top - d.sj] * fy;
            double vp = vis[below], vs = vis[below - d.sj];
            visf = vp * vs / (vp * (1.0 - fy) + vs * fy);
        }
        double term = flf * surface_tension_gradient(temp[top], a_oxygen, dgdtp) * grad / (visf * dzpbinv[d.nkm1]);
        phi[top] = phi[below] + term;
    } else if (t < ntop + nk_b * nj_b) {
        const int r = t - ntop;
        const int k = b.kstat + r / nj_b, j = b.jstat + r % nj_b;
        phi[cix(d, b.istat, j, k)] = 0.0;
        phi[cix(d, b.iend, j, k)] = 0.0;
    }
    }
}
};
__global__ __launch_bounds__(256) void k_bc_flow(Dims d, BoxDev b0, int ivar, double a_oxygen, double dgdtp,
                                                 const double *__restrict__ dxpwinv,
                                                 const double *__restrict__ dypsinv,
                                                 const double *__restrict__ dzpbinv,
                                                 const double *__restrict__ fracx,
                                                 const double *__restrict__ fracy,
                                                 const double *__restrict__ temp,
                                                 const double *__restrict__ fracl,
                                                 const double *__restrict__ vis, double *phi, const int *dbox,
                                                 const int *skip) {
    BcFlowK::run(vb_hw(), d, b0, ivar, a_oxygen, dgdtp, dxpwinv, dypsinv, dzpbinv, fracx, fracy, temp, fracl, vis, phi, dbox, skip);
}

// k_bc_flow and the preparation of the box passes' buffers (BoxPrepK) in
// one launch: they touch different memory, and both end before the box
// passes start
struct BcFlowPrepK {
    static constexpr int family = 23, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BoxDev b0, int ivar, double a_oxygen, double dgdtp,
                                               const double *__restrict__ dxpwinv,
                                               const double *__restrict__ dypsinv,
                                               const double *__restrict__ dzpbinv,
                                               const double *__restrict__ fracx,
                                               const double *__restrict__ fracy,
                                               const double *__restrict__ temp,
                                               const double *__restrict__ fracl,
                                               const double *__restrict__ vis, double *phi, const int *dbox,
                                               const int *skip, double *bufs, size_t bstride, int passes,
                                               unsigned int *claim) {
        BoxPrepK::run(vb, d, bufs, bstride, passes, d.nkm2, claim, skip, dbox);
        BcFlowK::run(vb, d, b0, ivar, a_oxygen, dgdtp, dxpwinv, dypsinv, dzpbinv, fracx, fracy, temp, fracl, vis, phi,
                     dbox, skip);
    }
};
__global__ __launch_bounds__(256) void k_bc_flow_prep(Dims d, BoxDev b0, int ivar, double a_oxygen, double dgdtp,
                                                      const double *__restrict__ dxpwinv,
                                                      const double *__restrict__ dypsinv,
                                                      const double *__restrict__ dzpbinv,
                                                      const double *__restrict__ fracx,
                                                      const double *__restrict__ fracy,
                                                      const double *__restrict__ temp,
                                                      const double *__restrict__ fracl,
                                                      const double *__restrict__ vis, double *phi, const int *dbox,
                                                      const int *skip, double *bufs, size_t bstride, int passes,
                                                      unsigned int *claim) {
    BcFlowPrepK::run(vb_hw(), d, b0, ivar, a_oxygen, dgdtp, dxpwinv, dypsinv, dzpbinv, fracx, fracy, temp, fracl, vis,
                     phi, dbox, skip, bufs, bstride, passes, claim);
}

struct MomConst {
    double delt, tsolid, urf, darcy, grav, beta, scanvel;
    int scheme, steady;
};

// Coefficients, sources, relaxation, solid pinning and residual for one
// velocity component on the staggered grid:
//   discretizeEquations(1/2/3) CFDSolverManager.cpp:771-1111
//   calculateSource(1/2/3)     CFDSolverManager.cpp:1408-1644
//   calculateResiduals(1/2/3)  CFDSolverManager.cpp:1960-2046
// IVAR selects the staggering: 1 = u (x faces), 2 = v (y faces), 3 = w (z faces).
template <int IVAR>
struct MomentumK {
    static constexpr int family = 16 + IVAR, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BoxDev b0, MomConst q, DevGeom g,
                                                     const double *__restrict__ U, const double *__restrict__ Vv,
                                                     const double *__restrict__ W, const double *__restrict__ V,
                                                     const double *__restrict__ rho,
                                                     const double *__restrict__ pressure,
                                                     const double *__restrict__ phinot,
                                                     const double *__restrict__ fracl,
                                                     const double *__restrict__ temp, double *an_o, double *as_o,
                                                     double *ae_o, double *aw_o, double *at_o, double *ab_o,
                                                     double *ap_o, double *su_o, double *dco, double *partial,
                                                     const int *dbox, const int *skip) {
    __shared__ double scratch[2 * 4];
    if (skip && *skip) return;
    const BoxTiles bt = box_tiles(d, b0, dbox);
    const BoxDev b = bt.b;
    const int sch = q.scheme;
    double resid = 0.0, umax = 0.0;
    for (int tt = vb.lin(); tt < bt.n; tt += vb.count()) {
    const int i = b.istat + 1 + (tt % bt.gx) * TX + threadIdx.x;
    const int j = b.jstat + ((tt / bt.gx) % bt.gy) * TY + threadIdx.y;
    const int k = b.kstat + tt / (bt.gx * bt.gy);
    if (i <= b.iend - 1 && j <= b.jend) {
        const size_t c = cix(d, i, j, k);
        const size_t sj = d.sj, sk = d.sk;
        const double r = rho[c];
        const double *phi = (IVAR == 1) ? U : (IVAR == 2) ? Vv : W;
        double fn, fs, fe, fw, ft, fb, dn, ds, de, dw, dt, db, vol, visn, viss, vise, visw, vist, visb;
        double aik, ajk, aij, pgrad;
        if (IVAR == 1) {
            const double fx = g.fracx[i - 1];
            double vn = Vv[c + sj] * (1.0 - fx) + Vv[c + sj - 1] * fx;
            double vs = Vv[c] * (1.0 - fx) + Vv[c - 1] * fx;
            double ue = (U[c + 1] + U[c]) * 0.5, uw = (U[c - 1] + U[c]) * 0.5;
            double wt = W[c + sk] * (1.0 - fx) + W[c + sk - 1] * fx;
            double wb = W[c] * (1.0 - fx) + W[c - 1] * fx;
            aik = g.areauik[k * d.ni + i]; ajk = g.areajk[k * d.nj + j]; aij = g.areauij[j * d.ni + i];
            fn = vn * r * aik; fs = vs * r * aik; fe = ue * r * ajk; fw = uw * r * ajk; ft = wt * r * aij; fb = wb * r * aij;
            double visu = hmean(V[c], V[c - 1], fx, 1.0 - fx);
            double visun = hmean(V[c + sj], V[c + sj - 1], fx, 1.0 - fx);
            double visus = hmean(V[c - sj], V[c - sj - 1], fx, 1.0 - fx);
            visn = (j == d.njm2) ? visu : hmean(visu, visun, 1.0 - g.fracy[j], g.fracy[j]);
            viss = (j == 1) ? visu : hmean(visu, visus, g.fracy[j - 1], 1.0 - g.fracy[j - 1]);
            vise = V[c]; visw = V[c - 1];
            double visut = hmean(V[c + sk], V[c + sk - 1], fx, 1.0 - fx);
            double visub = hmean(V[c - sk], V[c - sk - 1], fx, 1.0 - fx);
            vist = (k == d.nkm2) ? visu : hmean(visut, visu, g.fracz[k], 1.0 - g.fracz[k]);
            visb = (k == 1) ? visu : hmean(visub, visu, 1.0 - g.fracz[k - 1], g.fracz[k - 1]);
            dn = visn * aik * g.dypsinv[j + 1];
            ds = viss * aik * g.dypsinv[j];
            de = vise * ajk / (g.xu[i + 1] - g.xu[i]);
            dw = visw * ajk / (g.xu[i] - g.xu[i - 1]);
            dt = vist * aij * g.dzpbinv[k + 1];
            db = visb * aij * g.dzpbinv[k];
            vol = (g.x[i] - g.x[i - 1]) * (g.yv[j + 1] - g.yv[j]) * (g.zw[k + 1] - g.zw[k]);
            pgrad = ajk * (pressure[c - 1] - pressure[c]);
            double dudxp = (U[c + 1] - U[c]) / (g.xu[i + 1] - g.xu[i]);
            double dudxm = (U[c] - U[c - 1]) / (g.xu[i] - g.xu[i - 1]);
            double dvdxp = (Vv[c + sj] - Vv[c + sj - 1]) * g.dxpwinv[i];
            double dvdxm = (Vv[c] - Vv[c - 1]) * g.dxpwinv[i];
            double dwdxp = (W[c + sk] - W[c + sk - 1]) * g.dxpwinv[i];
            double dwdxm = (W[c] - W[c - 1]) * g.dxpwinv[i];
            // cross-derivative viscous terms, added in the reference's order
            double t1 = (vise * dudxp - visw * dudxm) * ajk;
            double t2 = (visn * dvdxp - viss * dvdxm) * aik;
            double t3 = (vist * dwdxp - visb * dwdxm) * aij;
            double an = coef_minus(sch, dn, fn), as = coef_plus(sch, ds, fs);
            double ae = coef_minus(sch, de, fe), aw = coef_plus(sch, dw, fw);
            double at = coef_minus(sch, dt, ft), ab = coef_plus(sch, db, fb);
            double apnot = r * vol / q.delt;
            double delf = fn - fs + fe - fw + ft - fb;
            double cp0 = dmax(0.0, delf), cp1 = dmin(0.0, delf);
            double su = -cp1 * U[c];
            su += pgrad;
            double sp = -cp0;
            su += apnot * phinot[c];
            su += t1;
            su += t2;
            su += t3;
            // Darcy damping in the mushy zone (+ scanning term in steady state)
            const double flf = fracl[c] * (1.0 - fx) + fracl[c - 1] * fx;
            if (flf > 0) {
                double term = q.darcy * ((1.0 - flf) * (1.0 - flf)) / ((flf * flf * flf) + AMCFD_SMALL);
                sp -= term * vol;
                if (q.steady) su += r * q.scanvel * ajk * (U[c - 1] - U[c]);
            }
            if (k == d.nkm2) { su += at * U[c + sk]; sp -= at; at = 0.0; }
            double ap = an + as + ae + aw + at + ab + apnot - sp;
            double dd = ajk / ap;
            ap /= q.urf;
            su += (1.0 - q.urf) * ap * U[c];
            dd *= q.urf;
            if (dmin(temp[c], temp[c - 1]) <= q.tsolid) {
                su = 0.0; an = 0.0; as = 0.0; ae = 0.0; aw = 0.0; at = 0.0; ab = 0.0; ap = AMCFD_LARGE;
            }
            an_o[c] = an; as_o[c] = as; ae_o[c] = ae; aw_o[c] = aw; at_o[c] = at; ab_o[c] = ab;
            ap_o[c] = ap; su_o[c] = su; dco[c] = dd;
            double res = an * phi[c + sj] + as * phi[c - sj] + ae * phi[c + 1] + aw * phi[c - 1] + at * phi[c + sk] +
                         ab * phi[c - sk] + su - ap * phi[c];
            resid += fabs(res);
            // first tile: plain assignment, as with one tile per workgroup
            if (k == d.nkm2) umax = (tt == vb.lin()) ? fabs(U[c + sk]) : dmax(umax, fabs(U[c + sk]));
        } else if (IVAR == 2) {
            const double fy = g.fracy[j - 1];
            double vn = (Vv[c] + Vv[c + sj]) * 0.5, vs = (Vv[c] + Vv[c - sj]) * 0.5;
            double ue = U[c + 1] * (1.0 - fy) + U[c - sj + 1] * fy;
            double uw = U[c] * (1.0 - fy) + U[c - sj] * fy;
            double wt = W[c + sk] * (1.0 - fy) + W[c + sk - sj] * fy;
            double wb = W[c] * (1.0 - fy) + W[c - sj] * fy;
            aik = g.areaik[k * d.ni + i]; ajk = g.areavjk[k * d.nj + j]; aij = g.areavij[j * d.ni + i];
            fn = vn * r * aik; fs = vs * r * aik; fe = ue * r * ajk; fw = uw * r * ajk; ft = wt * r * aij; fb = wb * r * aij;
            visn = V[c]; viss = V[c - sj];
            double visv = hmean(V[c], V[c - sj], fy, 1.0 - fy);
            double visve = hmean(V[c + 1], V[c - sj + 1], fy, 1.0 - fy);
            double visvw = hmean(V[c - 1], V[c - sj - 1], fy, 1.0 - fy);
            vise = (i == d.nim2) ? visv : hmean(visv, visve, 1.0 - g.fracx[i], g.fracx[i]);
            visw = (i == 1) ? visv : hmean(visv, visvw, g.fracx[i - 1], 1.0 - g.fracx[i - 1]);
            double visvt = hmean(V[c + sk], V[c + sk - sj], fy, 1.0 - fy);
            double visvb = hmean(V[c - sk], V[c - sk - sj], fy, 1.0 - fy);
            vist = (k == d.nkm2) ? visv : hmean(visv, visvt, 1.0 - g.fracz[k], g.fracz[k]);
            visb = (k == 1) ? visv : hmean(visv, visvb, g.fracz[k - 1], 1.0 - g.fracz[k - 1]);
            dn = visn * aik / (g.yv[j + 1] - g.yv[j]);
            ds = viss * aik / (g.yv[j] - g.yv[j - 1]);
            de = vise * ajk * g.dxpwinv[i + 1];
            dw = visw * ajk * g.dxpwinv[i];
            dt = vist * aij * g.dzpbinv[k + 1];
            db = visb * aij * g.dzpbinv[k];
            vol = (g.xu[i + 1] - g.xu[i]) * (g.y[j] - g.y[j - 1]) * (g.zw[k + 1] - g.zw[k]);
            pgrad = aik * (pressure[c - sj] - pressure[c]);
            double dudyp = (U[c + 1] - U[c - sj + 1]) * g.dypsinv[j];
            double dudym = (U[c] - U[c - sj]) * g.dypsinv[j];
            double dvdyp = (Vv[c + sj] - Vv[c]) / (g.yv[j + 1] - g.yv[j]);
            double dvdym = (Vv[c] - Vv[c - sj]) / (g.yv[j] - g.yv[j - 1]);
            double dwdyp = (W[c + sk] - W[c + sk - sj]) * g.dypsinv[j];
            double dwdym = (W[c] - W[c - sj]) * g.dypsinv[j];
            double t1 = (vise * dudyp - visw * dudym) * ajk;
            double t2 = (visn * dvdyp - viss * dvdym) * aik;
            double t3 = (vist * dwdyp - visb * dwdym) * aij;
            double an = coef_minus(sch, dn, fn), as = coef_plus(sch, ds, fs);
            double ae = coef_minus(sch, de, fe), aw = coef_plus(sch, dw, fw);
            double at = coef_minus(sch, dt, ft), ab = coef_plus(sch, db, fb);
            double apnot = r * vol / q.delt;
            double delf = fn - fs + fe - fw + ft - fb;
            double cp0 = dmax(0.0, delf), cp1 = dmin(0.0, delf);
            double su = -cp1 * Vv[c];
            su += pgrad;
            double sp = -cp0;
            su += apnot * phinot[c];
            su += t1;
            su += t2;
            su += t3;
            const double flf = fracl[c] * (1.0 - fy) + fracl[c - sj] * fy;
            if (flf > 0) {
                double term = q.darcy * ((1.0 - flf) * (1.0 - flf)) / ((flf * flf * flf) + AMCFD_SMALL);
                sp -= term * vol;
                if (q.steady) su += r * q.scanvel * ajk * (Vv[c - 1] - Vv[c]);
            }
            if (k == d.nkm2) { su += at * Vv[c + sk]; sp -= at; at = 0.0; }
            double ap = an + as + ae + aw + at + ab + apnot - sp;
            double dd = aik / ap;
            ap /= q.urf;
            su += (1.0 - q.urf) * ap * Vv[c];
            dd *= q.urf;
            if (dmin(temp[c], temp[c - sj]) <= q.tsolid) {
                su = 0.0; an = 0.0; as = 0.0; ae = 0.0; aw = 0.0; at = 0.0; ab = 0.0; ap = AMCFD_LARGE;
            }
            an_o[c] = an; as_o[c] = as; ae_o[c] = ae; aw_o[c] = aw; at_o[c] = at; ab_o[c] = ab;
            ap_o[c] = ap; su_o[c] = su; dco[c] = dd;
            double res = an * phi[c + sj] + as * phi[c - sj] + ae * phi[c + 1] + aw * phi[c - 1] + at * phi[c + sk] +
                         ab * phi[c - sk] + su - ap * phi[c];
            resid += fabs(res);
        } else {
            const double fz = g.fracz[k - 1];
            double vn = Vv[c + sj] * (1.0 - fz) + Vv[c - sk + sj] * fz;
            double vs = Vv[c] * (1.0 - fz) + Vv[c - sk] * fz;
            double ue = U[c + 1] * (1.0 - fz) + U[c - sk + 1] * fz;
            double uw = U[c] * (1.0 - fz) + U[c - sk] * fz;
            double wt = (W[c] + W[c + sk]) * 0.5, wb = (W[c] + W[c - sk]) * 0.5;
            aik = g.areawik[k * d.ni + i]; ajk = g.areawjk[k * d.nj + j]; aij = g.areaij[j * d.ni + i];
            fn = vn * r * aik; fs = vs * r * aik; fe = ue * r * ajk; fw = uw * r * ajk; ft = wt * r * aij; fb = wb * r * aij;
            double visw0 = hmean(V[c], V[c - sk], fz, 1.0 - fz);
            double viswn = hmean(V[c + sj], V[c - sk + sj], fz, 1.0 - fz);
            double visws = hmean(V[c - sj], V[c - sk - sj], fz, 1.0 - fz);
            visn = (j == d.njm2) ? visw0 : hmean(visw0, viswn, 1.0 - g.fracy[j], g.fracy[j]);
            viss = (j == 1) ? visw0 : hmean(visw0, visws, g.fracy[j - 1], 1.0 - g.fracy[j - 1]);
            double viswe = hmean(V[c + 1], V[c - sk + 1], fz, 1.0 - fz);
            double visww = hmean(V[c - 1], V[c - sk - 1], fz, 1.0 - fz);
            vise = (i == d.nim2) ? visw0 : hmean(visw0, viswe, 1.0 - g.fracx[i], g.fracx[i]);
            visw = (i == 1) ? visw0 : hmean(visw0, visww, g.fracx[i - 1], 1.0 - g.fracx[i - 1]);
            vist = V[c]; visb = V[c - sk];
            dn = visn * aik * g.dypsinv[j + 1];
            ds = viss * aik * g.dypsinv[j];
            de = vise * ajk * g.dxpwinv[i + 1];
            dw = visw * ajk * g.dxpwinv[i];
            dt = vist * aij / (g.zw[k + 1] - g.zw[k]);
            db = visb * aij / (g.zw[k] - g.zw[k - 1]);
            vol = (g.xu[i + 1] - g.xu[i]) * (g.yv[j + 1] - g.yv[j]) * (g.z[k] - g.z[k - 1]);
            pgrad = aij * (pressure[c - sk] - pressure[c]);
            double dudzp = (U[c + 1] - U[c - sk + 1]) * g.dzpbinv[k];
            double dudzm = (U[c] - U[c - sk]) * g.dzpbinv[k];
            double dvdzp = (Vv[c + sj] - Vv[c - sk + sj]) * g.dzpbinv[k];
            double dvdzm = (Vv[c] - Vv[c - sk]) * g.dzpbinv[k];
            double dwdzp = (W[c + sk] - W[c]) / (g.zw[k + 1] - g.zw[k]);
            double dwdzm = (W[c] - W[c - sk]) / (g.zw[k] - g.zw[k - 1]);
            double t1 = (vise * dudzp - visw * dudzm) * ajk;
            double t2 = (visn * dvdzp - viss * dvdzm) * aik;
            double t3 = (vist * dwdzp - visb * dwdzm) * aij;
            double an = coef_minus(sch, dn, fn), as = coef_plus(sch, ds, fs);
            double ae = coef_minus(sch, de, fe), aw = coef_plus(sch, dw, fw);
            double at = coef_minus(sch, dt, ft), ab = coef_plus(sch, db, fb);
            double apnot = r * vol / q.delt;
            double delf = fn - fs + fe - fw + ft - fb;
            double cp0 = dmax(0.0, delf), cp1 = dmin(0.0, delf);
            double su = -cp1 * W[c];
            su += pgrad;
            double sp = -cp0;
            su += apnot * phinot[c];
            su += t1;
            su += t2;
            su += t3;
            const double flf = fracl[c] * (1.0 - fz) + fracl[c - sk] * fz;
            const double boufac = r * q.grav * q.beta;
            if (flf > 0) {
                double term = q.darcy * ((1.0 - flf) * (1.0 - flf)) / ((flf * flf * flf) + AMCFD_SMALL);
                sp -= term * vol;
                if (q.steady) su += r * q.scanvel * ajk * (W[c - 1] - W[c]);
                double tw = temp[c] * (1.0 - fz) + temp[c - sk] * fz;
                su += boufac * vol * (tw - q.tsolid);
            }
            double ap = an + as + ae + aw + at + ab + apnot - sp;
            double dd = aij / ap;
            ap /= q.urf;
            su += (1.0 - q.urf) * ap * W[c];
            dd *= q.urf;
            if (dmin(temp[c], temp[c - sk]) <= q.tsolid) {
                su = 0.0; an = 0.0; as = 0.0; ae = 0.0; aw = 0.0; at = 0.0; ab = 0.0; ap = AMCFD_LARGE;
            }
            an_o[c] = an; as_o[c] = as; ae_o[c] = ae; aw_o[c] = aw; at_o[c] = at; ab_o[c] = ab;
            ap_o[c] = ap; su_o[c] = su; dco[c] = dd;
            double res = an * phi[c + sj] + as * phi[c - sj] + ae * phi[c + 1] + aw * phi[c - 1] + at * phi[c + sk] +
                         ab * phi[c - sk] + su - ap * phi[c];
            resid += fabs(res);
        }
    }
    }
    // block partials: [sum |resid|, max |u_top|]
    const int lane = threadIdx.x + threadIdx.y * TX;
    resid = wave_sum(resid);
    umax = wave_max(umax);
    if ((lane & 63) == 0) {
        scratch[lane >> 6] = resid;
        scratch[4 + (lane >> 6)] = umax;
    }
    __syncthreads();
    if (lane == 0) {
        const int bid = vb.x + vb.gx * (vb.y + vb.gy * vb.z);
        double s = 0.0, m = 0.0;
        for (int w = 0; w < 4; ++w) { s += scratch[w]; m = dmax(m, scratch[4 + w]); }
        partial[(size_t)bid * 3] = s;
        partial[(size_t)bid * 3 + 1] = m;
        partial[(size_t)bid * 3 + 2] = 0.0;
    }
}
};
template <int IVAR>
__global__ __launch_bounds__(TX *TY) void k_momentum(Dims d, BoxDev b0, MomConst q, DevGeom g,
                                                     const double *__restrict__ U, const double *__restrict__ Vv,
                                                     const double *__restrict__ W, const double *__restrict__ V,
                                                     const double *__restrict__ rho,
                                                     const double *__restrict__ pressure,
                                                     const double *__restrict__ phinot,
                                                     const double *__restrict__ fracl,
                                                     const double *__restrict__ temp, double *an_o, double *as_o,
                                                     double *ae_o, double *aw_o, double *at_o, double *ab_o,
                                                     double *ap_o, double *su_o, double *dco, double *partial,
                                                     const int *dbox, const int *skip) {
    MomentumK<IVAR>::run(vb_hw(), d, b0, q, g, U, Vv, W, V, rho, pressure, phinot, fracl, temp, an_o, as_o, ae_o, aw_o, at_o, ab_o, ap_o, su_o, dco, partial, dbox, skip);
}

// pressure-correction coefficients, mass imbalance, solid pinning and the
// continuity normaliser: CFDSolverManager.cpp:1112-1156, 1645-1676, 2047-2071
struct PressureK {
    static constexpr int family = 20, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BoxDev b0, double tsolid, DevGeom g,
                                                     const double *__restrict__ U, const double *__restrict__ Vv,
                                                     const double *__restrict__ W, const double *__restrict__ rho,
                                                     const double *__restrict__ dux, const double *__restrict__ dvy,
                                                     const double *__restrict__ dwz,
                                                     const double *__restrict__ temp, double *an_o, double *as_o,
                                                     double *ae_o, double *aw_o, double *at_o, double *ab_o,
                                                     double *ap_o, double *su_o, double *partial, const int *dbox,
                                                     const int *skip) {
    __shared__ double scratch[2 * 4];
    if (skip && *skip) return;
    const BoxTiles bt = box_tiles(d, b0, dbox);
    const BoxDev b = bt.b;
    double dsum = 0.0, msum = 0.0;
    for (int tt = vb.lin(); tt < bt.n; tt += vb.count()) {
    const int i = b.istat + 1 + (tt % bt.gx) * TX + threadIdx.x;
    const int j = b.jstat + ((tt / bt.gx) % bt.gy) * TY + threadIdx.y;
    const int k = b.kstat + tt / (bt.gx * bt.gy);
    if (i <= b.iend - 1 && j <= b.jend) {
        const size_t c = cix(d, i, j, k);
        const size_t sj = d.sj, sk = d.sk;
        const double r = rho[c];
        const double aik = g.areaik[k * d.ni + i], ajk = g.areajk[k * d.nj + j], aij = g.areaij[j * d.ni + i];
        double an = aik * dvy[c + sj] * r, as = aik * dvy[c] * r;
        double ae = ajk * dux[c + 1] * r, aw = ajk * dux[c] * r;
        double at = aij * dwz[c + sk] * r, ab = aij * dwz[c] * r;
        double fn = Vv[c + sj] * aik * r, fs = Vv[c] * aik * r;
        double fe = U[c + 1] * ajk * r, fw = U[c] * ajk * r;
        double ft = W[c + sk] * aij * r, fb = W[c] * aij * r;
        double delf = fn - fs + fe - fw + ft - fb;
        double su = -delf;
        msum += fabs(delf);
        double ap = an + as + ae + aw + at + ab - 0.0;
        if (temp[c] <= tsolid) {
            su = 0.0; an = 0.0; as = 0.0; ae = 0.0; aw = 0.0; at = 0.0; ab = 0.0; ap = AMCFD_LARGE;
        }
        an_o[c] = an; as_o[c] = as; ae_o[c] = ae; aw_o[c] = aw; at_o[c] = at; ab_o[c] = ab;
        ap_o[c] = ap; su_o[c] = su;
        double dtp = (fabs(U[c]) + fabs(U[c + 1])) * ajk + (fabs(Vv[c]) + fabs(Vv[c + sj])) * aik +
                     (fabs(W[c]) + fabs(W[c + sk])) * aij;
        dsum += 0.5 * fabs(dtp);
    }
    }
    const int lane = threadIdx.x + threadIdx.y * TX;
    dsum = wave_sum(dsum);
    msum = wave_sum(msum);
    if ((lane & 63) == 0) {
        scratch[lane >> 6] = dsum;
        scratch[4 + (lane >> 6)] = msum;
    }
    __syncthreads();
    if (lane == 0) {
        const int bid = vb.x + vb.gx * (vb.y + vb.gy * vb.z);
        double s = 0.0, m = 0.0;
        for (int w = 0; w < 4; ++w) { s += scratch[w]; m += scratch[4 + w]; }
        partial[(size_t)bid * 3] = 0.0;
        partial[(size_t)bid * 3 + 1] = s;
        partial[(size_t)bid * 3 + 2] = m;
    }
}
};
__global__ __launch_bounds__(TX *TY) void k_pressure(Dims d, BoxDev b0, double tsolid, DevGeom g,
                                                     const double *__restrict__ U, const double *__restrict__ Vv,
                                                     const double *__restrict__ W, const double *__restrict__ rho,
                                                     const double *__restrict__ dux, const double *__restrict__ dvy,
                                                     const double *__restrict__ dwz,
                                                     const double *__restrict__ temp, double *an_o, double *as_o,
                                                     double *ae_o, double *aw_o, double *at_o, double *ab_o,
                                                     double *ap_o, double *su_o, double *partial, const int *dbox,
                                                     const int *skip) {
    PressureK::run(vb_hw(), d, b0, tsolid, g, U, Vv, W, rho, dux, dvy, dwz, temp, an_o, as_o, ae_o, aw_o, at_o, ab_o, ap_o, su_o, partial, dbox, skip);
}

// velocity correction from p', CFDSolverManager.cpp:2719-2742
struct CorrectVelK {
    static constexpr int family = 21, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BoxDev b0, double ts, const double *__restrict__ temp,
                                                        const double *__restrict__ pp,
                                                        const double *__restrict__ dux,
                                                        const double *__restrict__ dvy,
                                                        const double *__restrict__ dwz, double *u, double *v,
                                                        double *w, const int *dbox, const int *skip) {
    if (skip && *skip) return;
    const BoxTiles bt = box_tiles(d, b0, dbox);
    const BoxDev b = bt.b;
    for (int tt = vb.lin(); tt < bt.n; tt += vb.count()) {
        const int i = b.istat + 1 + (tt % bt.gx) * TX + threadIdx.x;
        const int j = b.jstat + ((tt / bt.gx) % bt.gy) * TY + threadIdx.y;
        const int k = b.kstat + tt / (bt.gx * bt.gy);
        if (i > b.iend - 1 || j > b.jend) continue;
        const size_t c = cix(d, i, j, k);
        const double tP = temp[c];
        if (dmin(tP, temp[c - 1]) > ts) u[c] += dux[c] * (pp[c - 1] - pp[c]);
        if (dmin(tP, temp[c - d.sj]) > ts) v[c] += dvy[c] * (pp[c - d.sj] - pp[c]);
        if (dmin(tP, temp[c - d.sk]) > ts) w[c] += dwz[c] * (pp[c - d.sk] - pp[c]);
    }
}
};
__global__ __launch_bounds__(TX *TY) void k_correct_vel(Dims d, BoxDev b0, double ts, const double *__restrict__ temp,
                                                        const double *__restrict__ pp,
                                                        const double *__restrict__ dux,
                                                        const double *__restrict__ dvy,
                                                        const double *__restrict__ dwz, double *u, double *v,
                                                        double *w, const int *dbox, const int *skip) {
    CorrectVelK::run(vb_hw(), d, b0, ts, temp, pp, dux, dvy, dwz, u, v, w, dbox, skip);
}

// pressure update, CFDSolverManager.cpp:2744-2757
struct CorrectPK {
    static constexpr int family = 22, rank = 0;
    static __device__ __forceinline__ void run(const VB &vb, Dims d, BoxDev b0, double ts, double urfp,
                                                      const double *__restrict__ temp, double *pp,
                                                      double *pressure, const int *dbox, const int *skip) {
    if (skip && *skip) return;
    const BoxTiles bt = box_tiles(d, b0, dbox);
    const BoxDev b = bt.b;
    for (int tt = vb.lin(); tt < bt.n; tt += vb.count()) {
        const int i = b.istat + 1 + (tt % bt.gx) * TX + threadIdx.x;
        const int j = b.jstat + ((tt / bt.gx) % bt.gy) * TY + threadIdx.y;
        const int k = b.kstat + tt / (bt.gx * bt.gy);
        if (i > b.iend - 1 || j > b.jend) continue;
        const size_t c = cix(d, i, j, k);
        if (temp[c] > ts) {
            pressure[c] += urfp * pp[c];
            pp[c] = 0.0;
        }
    }
}
};
__global__ __launch_bounds__(TX *TY) void k_correct_p(Dims d, BoxDev b0, double ts, double urfp,
                                                      const double *__restrict__ temp, double *pp,
                                                      double *pressure, const int *dbox, const int *skip) {
    CorrectPK::run(vb_hw(), d, b0, ts, urfp, temp, pp, pressure, dbox, skip);
}

dim3 box_grid(const Box &b, const Dims &d) {
    int ni_b = b.iend - b.istat - 1, nj_b = b.jend - b.jstat + 1, nk_b = d.nkm1 - b.kstat;
    if (ni_b < 1 || nj_b < 1 || nk_b < 1) return dim3(0, 0, 0);
    return dim3((ni_b + TX - 1) / TX, (nj_b + TY - 1) / TY, nk_b);
}

BoxDev to_dev(const Box &b) { return BoxDev{b.istat, b.iend, b.jstat, b.jend, b.kstat}; }

}  // namespace

int flow_partial_blocks(const Box &b, const Dims &d) {
    dim3 g = box_grid(b, d);
    return g.x * g.y * g.z;
}

// grid of a box launch: the box's own tiles (FP64 mode), or a fixed number
// of workgroups that stride over the device box (graphs)
static dim3 bl_grid(const Box &b, const Dims &d, const BoxLaunch &bl) {
    if (bl.dbox) return dim3(bl.nblocks, 1, 1);
    return box_grid(b, d);
}

void launch_clean_velocities(hipStream_t st, const Dims &d, DevFields f, const Box &b, double tsolid,
                             const BoxLaunch &bl) {
    dim3 g = bl_grid(b, d, bl);
    if (!g.x) return;
    dispatch<CleanK, TX * TY>(k_clean, g, dim3(TX, TY), 0, st, d, to_dev(b), tsolid, f.temp, f.uvel, f.vvel, f.wvel, bl.dbox,
                       bl.skip);
}

void launch_bc_flow(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, const Box &b,
                    const amcfd_params &p, int ivar, const BoxLaunch &bl, const BoxPrep *prep) {
    double *phi = ivar == 1 ? f.uvel : ivar == 2 ? f.vvel : ivar == 3 ? f.wvel : f.pp;
    int nb;
    if (bl.dbox) {
        nb = bl.nblocks;
    } else {
        int ni_b = b.iend - b.istat - 1, nj_b = b.jend - b.jstat + 1, nk_b = d.nkm1 - b.kstat;
        int n = ((ivar <= 2) ? ni_b * nj_b : 0) + nk_b * nj_b;
        if (n <= 0) return;
        nb = (n + 255) / 256;
    }
    if (prep && bl.dbox)
        dispatch<BcFlowPrepK, 256>(k_bc_flow_prep, dim3(nb), dim3(256), 0, st, d, to_dev(b), ivar, p.a_oxygen, p.dgdtp,
                                   g.dxpwinv, g.dypsinv, g.dzpbinv, g.fracx, g.fracy, f.temp, f.fracl, f.vis, phi, bl.dbox,
                                   bl.skip, prep->bufs, prep->bstride, prep->passes, prep->claim);
    else
        dispatch<BcFlowK, 256>(k_bc_flow, dim3(nb), dim3(256), 0, st, d, to_dev(b), ivar, p.a_oxygen, p.dgdtp,
                               g.dxpwinv, g.dypsinv, g.dzpbinv, g.fracx, g.fracy, f.temp, f.fracl, f.vis, phi, bl.dbox,
                               bl.skip);
}

void launch_momentum(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr, const Box &b,
                     const amcfd_params &p, double scanvel, int ivar, const BoxLaunch &bl) {
    dim3 gr = bl_grid(b, d, bl);
    if (!gr.x) return;
    double urf = ivar == 1 ? p.urfu : ivar == 2 ? p.urfv : p.urfw;
    MomConst q{p.delt, p.tsolid, urf, p.darcyKo, p.grav, p.beta, scanvel, p.discretization, p.steady};
    double *part = pr.flow + (size_t)(ivar - 1) * pr.nblk_f * 3;
    if (ivar == 1)
        dispatch<MomentumK<1>, TX * TY>(k_momentum<1>, gr, dim3(TX, TY), 0, st, d, to_dev(b), q, g, f.uvel, f.vvel, f.wvel, f.vis,
                           f.rho, f.pressure, f.unot, f.fracl, f.temp, f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su,
                           f.dux, part, bl.dbox, bl.skip);
    else if (ivar == 2)
        dispatch<MomentumK<2>, TX * TY>(k_momentum<2>, gr, dim3(TX, TY), 0, st, d, to_dev(b), q, g, f.uvel, f.vvel, f.wvel, f.vis,
                           f.rho, f.pressure, f.vnot, f.fracl, f.temp, f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su,
                           f.dvy, part, bl.dbox, bl.skip);
    else
        dispatch<MomentumK<3>, TX * TY>(k_momentum<3>, gr, dim3(TX, TY), 0, st, d, to_dev(b), q, g, f.uvel, f.vvel, f.wvel, f.vis,
                           f.rho, f.pressure, f.wnot, f.fracl, f.temp, f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su,
                           f.dwz, part, bl.dbox, bl.skip);
}

void launch_pressure(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr, const Box &b,
                     const amcfd_params &p, const BoxLaunch &bl) {
    dim3 gr = bl_grid(b, d, bl);
    if (!gr.x) return;
    double *part = pr.flow + (size_t)3 * pr.nblk_f * 3;
    dispatch<PressureK, TX * TY>(k_pressure, gr, dim3(TX, TY), 0, st, d, to_dev(b), p.tsolid, g, f.uvel, f.vvel, f.wvel, f.rho,
                       f.dux, f.dvy, f.dwz, f.temp, f.an, f.as, f.ae, f.aw, f.at, f.ab, f.ap, f.su, part, bl.dbox,
                       bl.skip);
}

void launch_correct(hipStream_t st, const Dims &d, DevFields f, const Box &b, const amcfd_params &p,
                    const BoxLaunch &bl) {
    dim3 gr = bl_grid(b, d, bl);
    if (!gr.x) return;
    dispatch<CorrectVelK, TX * TY>(k_correct_vel, gr, dim3(TX, TY), 0, st, d, to_dev(b), p.tsolid, f.temp, f.pp, f.dux, f.dvy,
                       f.dwz, f.uvel, f.vvel, f.wvel, bl.dbox, bl.skip);
    dispatch<CorrectPK, TX * TY>(k_correct_p, gr, dim3(TX, TY), 0, st, d, to_dev(b), p.tsolid, p.urfp, f.temp, f.pp,
                       f.pressure, bl.dbox, bl.skip);
}

}  // namespace amcfd
