// Energy-based free surface, evaluated at output times:
// FreeSurfaceManager::execute / getDeltaVol / ebal
// (FreeSurfaceManager.cpp:72-342).
//
// The surface shift fi(x, y) of the molten top cells balances surface
// tension, gravity and a Lagrange multiplier; the multiplier is found by
// bisection so that the shifted volume takes up the added droplet volume.
// Every energy iteration is a Jacobi sweep over the top plane that depends on
// the previous one, and the bisection depends on every sweep, so the whole
// search runs in one workgroup (k_fs_solve) with barriers between the
// sweeps; the top plane (ni*nj doubles per array) stays in L2.  The volume
// sums run in the reference's (j, i) order on one lane.  pow(v, 1.5) is the
// correctly rounded value (pow15.h), which the C library's pow misses by one
// unit in the last place in about one call in 1300.  Then k_fs_profile
// re-interpolates node heights and temperatures along z for every column.

#include "fv.h"
#include "pow15.h"

namespace amcfd {

constexpr int FS_NT = 1024;

struct FsConst {
    double alam1, alam2, tolb, c1, gamma, urfree, tole, tsolid, vd;
    int nitb, nite;
};

namespace {

// one surface-shift solve for multiplier alam (FreeSurfaceManager::ebal),
// then its volume change (getDeltaVol); called by every thread of the block
__device__ double fs_delta_vol(const Dims &d, const FsConst &q, double alam, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ xu,
                               const double *__restrict__ yv, const double *__restrict__ areaij,
                               const double *__restrict__ temp, double *fi, double *zzm, double *s_red,
                               double *s_val, int *s_flag) {
    const int tid = threadIdx.x, ni = d.ni;
    const int n2 = d.ni * d.nj, ncell = d.nim2 * d.njm2;
    const size_t top = (size_t)d.nkm1 * d.sk;
    for (int c = tid; c < n2; c += FS_NT) zzm[c] = 0.0;
    __syncthreads();
    double difmax = 0.0;   // meaningful on thread 0
    for (int niter = 0; niter < q.nite;) {
        niter++;
        double pmax = -__builtin_inf();
        for (int t = tid; t < ncell; t += FS_NT) {
            const int i = 1 + t % d.nim2, j = 1 + t / d.nim2;
            const int c = j * ni + i;
            if (temp[top + c] < q.tsolid) continue;
            const double dely = yv[j + 1] - yv[j - 1], dely1 = y[j + 1] - y[j - 1];
            const double delyp = y[j + 1] - y[j], delym = y[j] - y[j - 1];
            const double delx = xu[i + 1] - xu[i - 1], delx1 = x[i + 1] - x[i - 1];
            const double delxp = x[i + 1] - x[i], delxm = x[i] - x[i - 1];
            const double c2 = 0.0 + alam;   // applied pressure is zero (:103-110)
            const double dzdx = (zzm[c + 1] - zzm[c - 1]) / delx1;
            const double dzdy = (zzm[c + ni] - zzm[c - ni]) / dely1;
            const double dzdyp1 = (zzm[c + ni + 1] - zzm[c - ni + 1]) / dely1;
            const double dzdym1 = (zzm[c + ni - 1] - zzm[c - ni - 1]) / dely1;
            const double d2zdxdy = (dzdyp1 - dzdym1) / delx1;
            const double c3 = pow15(1.0 + dzdx * dzdx + dzdy * dzdy);
            const double c4 = 1.0 + dzdy * dzdy, c5 = 1.0 + dzdx * dzdx;
            const double c6 = 2.0 * dzdx * dzdy * d2zdxdy;
            const double c7 = zzm[c + 1] / (delxp * delx) + zzm[c - 1] / (delxm * delx);
            const double c8 = zzm[c + ni] / (delyp * dely) + zzm[c - ni] / (delym * dely);
            const double c9 = 1.0 / (delxp * delx) + 1.0 / (delxm * delx);
            const double c10 = 1.0 / (delyp * dely) + 1.0 / (delym * dely);
            const double c11 = q.gamma / c3;
            const double term1 = c2 + c11 * (c4 * c7 + c5 * c8 - c6);
            const double term2 = q.c1 + c11 * (c4 * c9 + c5 * c10);
            const double f = q.urfree * term1 / term2 + (1.0 - q.urfree) * zzm[c];
            fi[c] = f;
            const double diff1 = fabs((f - zzm[c]) / zzm[c]);
            pmax = dmax(diff1, pmax);
        }
        pmax = wave_max(pmax);
        if ((tid & 63) == 0) s_red[tid >> 6] = pmax;
        __syncthreads();
        if (tid == 0) {
            double m = s_red[0];
            for (int w = 1; w < FS_NT / 64; ++w) m = dmax(m, s_red[w]);
            difmax = (m > difmax) ? m : difmax;
            *s_flag = difmax < q.tole;
        }
        __syncthreads();
        const int stop = *s_flag;
        __syncthreads();
        if (stop) break;
        for (int c = tid; c < n2; c += FS_NT) zzm[c] = fi[c];
        __syncthreads();
    }
    if (tid == 0) {
        double delv = 0.0;
        for (int j = 1; j <= d.njm2; ++j)
            for (int i = 1; i <= d.nim2; ++i) delv += areaij[j * ni + i] * fi[j * ni + i];
        *s_val = delv;
    }
    __syncthreads();
    const double r = *s_val;
    __syncthreads();
    return r;
}

}  // namespace

// the bisection of FreeSurfaceManager::execute (:119-166); out[0..5] =
// multiplier, last residual, min fi, max fi (first extremum, as
// std::min/max_element), bisection count, bracket-warning flag
__global__ __launch_bounds__(FS_NT) void k_fs_solve(Dims d, FsConst q, const double *__restrict__ x,
                                                    const double *__restrict__ y, const double *__restrict__ xu,
                                                    const double *__restrict__ yv,
                                                    const double *__restrict__ areaij,
                                                    const double *__restrict__ temp, double *fi, double *zzm,
                                                    double *out) {
    __shared__ double s_red[FS_NT / 64];
    __shared__ double s_val;
    __shared__ int s_flag;
    const double f = fs_delta_vol(d, q, q.alam1, x, y, xu, yv, areaij, temp, fi, zzm, s_red, &s_val, &s_flag) + q.vd;
    double fmid = fs_delta_vol(d, q, q.alam2, x, y, xu, yv, areaij, temp, fi, zzm, s_red, &s_val, &s_flag) + q.vd;
    const int warn = (f * fmid >= 0.0 && f != q.vd);
    double alamopt, dx, xmid = 0.0;
    if (f < 0.0) {
        alamopt = q.alam1;
        dx = q.alam2 - q.alam1;
    } else {
        alamopt = q.alam2;
        dx = q.alam1 - q.alam2;
    }
    int niter;
    for (niter = 0; niter < q.nitb; niter++) {
        dx *= 0.5;
        xmid = alamopt + dx;
        fmid = fs_delta_vol(d, q, xmid, x, y, xu, yv, areaij, temp, fi, zzm, s_red, &s_val, &s_flag) + q.vd;
        if (fmid < 0.0) alamopt = xmid;
        if (fabs(dx) < 1.0e-5 || (fabs(fmid) < q.tolb && niter != 0)) break;
    }
    if (threadIdx.x == 0) {
        const int n2 = d.ni * d.nj;
        double lo = fi[0], hi = fi[0];
        for (int c = 1; c < n2; ++c) {
            if (fi[c] < lo) lo = fi[c];
            if (hi < fi[c]) hi = fi[c];
        }
        out[0] = xmid;
        out[1] = fmid;
        out[2] = lo;
        out[3] = hi;
        out[4] = (double)niter;
        out[5] = (double)warn;
    }
}

// shifted node heights and temperatures, :169-194 (active extent)
__global__ __launch_bounds__(256) void k_fs_profile(Dims d, const double *__restrict__ z, const double *__restrict__ fi,
                                                    const double *__restrict__ temp, double *zr, double *tempr) {
    const size_t n = (size_t)d.nk * d.sk;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int k = (int)(c / d.sk);
    const size_t c2 = c - (size_t)k * d.sk;
    const double ztop = z[d.nkm1];
    const double factor = fi[c2] / ztop;
    const double h = z[k] * (1.0 - factor);
    zr[c] = h;
    if (h > ztop) return;
    if (h == ztop) tempr[c] = temp[(size_t)d.nkm1 * d.sk + c2];
    for (int l = d.nkm1 - 1; l >= 0; l--)
        if (z[l] <= h) {
            const double tl = temp[(size_t)l * d.sk + c2], tu = temp[(size_t)(l + 1) * d.sk + c2];
            tempr[c] = (h - z[l]) / (z[l + 1] - z[l]) * (tu - tl) + tl;
            break;
        }
}

void launch_free_surface(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, const amcfd_params &p,
                         double vd, double *out) {
    FsConst q{p.initlam1, p.initlam2, p.tolbisect, p.dens * p.grav, p.surftens, p.urfree, p.tolenergy, p.tsolid,
              vd, (int)p.maxiterbisect, (int)p.maxiterenergy};
    hipLaunchKernelGGL(k_fs_solve, dim3(1), dim3(FS_NT), 0, st, d, q, g.x, g.y, g.xu, g.yv, g.areaij, f.temp, f.fi,
                       f.zzm, out);
    const size_t n = (size_t)d.nk * d.sk;
    hipLaunchKernelGGL(k_fs_profile, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, g.z, f.fi, f.temp, f.zr,
                       f.tempr);
}

}  // namespace amcfd
