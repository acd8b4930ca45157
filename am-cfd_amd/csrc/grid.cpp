// Host-side set-up of the structured staggered grid and the initial state.
//
//   zone-wise power-law face placement   CFDSolverManager.cpp:128-220
//   inverse node spacings / face fractions CFDSolverManager.cpp:223-240
//   face areas                            CFDSolverManager.cpp:258-286
//   initial fields                        CFDSolverManager.cpp:429-525
//
// The grid is computed on the host in FP64 with the same expressions (so it
// is bitwise the reference grid) and uploaded: once, or every step when the
// mesh moves; cell volumes are recomputed on the fly in the kernels from the
// 1-D face coordinates.

#include "host.h"

#include <cmath>

namespace amcfd {

// Faces of every zone are placed; nodes only up to n-1 (the active extent):
// with a powder bed the reference regenerates the grid with its current nk,
// leaving the entries above it as they were.
static void place_axis(int nzones, const double *len, const double *ncv, const double *expo, int n,
                       std::vector<double> &face, std::vector<double> &node) {
    face[0] = 0.0;
    face[1] = 0.0;
    int base = 2;
    double origin = 0.0;
    for (int z = 0; z < nzones; ++z) {
        for (int c = 0; c < ncv[z]; ++c) {
            const double s = (double)(c + 1) / ncv[z];
            // expo >= 0: fine -> coarse along +axis; expo < 0: coarse -> fine
            const double t = (expo[z] >= 0.0) ? std::pow(s, expo[z]) : 1.0 - std::pow(1.0 - s, -expo[z]);
            face[base + c] = origin + len[z] * t;
        }
        base += (int)ncv[z];
        origin += len[z];
    }
    for (int i = 0; i < n - 1; ++i) node[i] = (face[i + 1] + face[i]) * 0.5;
    node[n - 1] = face[n - 1];
}

void HostGrid::build(const amcfd_params &p) {
    ni = p.ni;
    nj = p.nj;
    nk = p.nk;
    for (auto *v : {&x, &xu, &dxpwinv, &fracx}) v->assign(ni + 2, 0.0);
    for (auto *v : {&y, &yv, &dypsinv, &fracy}) v->assign(nj + 2, 0.0);
    for (auto *v : {&z, &zw, &dzpbinv, &fracz}) v->assign(nk + 2, 0.0);
    for (auto *v : {&areaij, &areauij, &areavij}) v->assign((size_t)nj * ni, 0.0);
    for (auto *v : {&areaik, &areawik, &areauik}) v->assign((size_t)nk * ni, 0.0);
    for (auto *v : {&areajk, &areavjk, &areawjk}) v->assign((size_t)nk * nj, 0.0);
    regenerate(p, nk);
}

void HostGrid::regenerate(const amcfd_params &p, int nka) {
    place_axis(p.nzx, p.xzone, p.ncvx, p.powrx, ni, xu, x);
    place_axis(p.nzy, p.yzone, p.ncvy, p.powry, nj, yv, y);
    place_axis(p.nzz, p.zzone, p.ncvz, p.powrz, nka, zw, z);
    auto inv = [](const std::vector<double> &c, int n, std::vector<double> &out) {
        for (int i = 1; i < n; ++i) out[i] = 1.0 / (c[i] - c[i - 1]);
    };
    inv(x, ni, dxpwinv);
    inv(y, nj, dypsinv);
    inv(z, nka, dzpbinv);
    auto frac = [](const std::vector<double> &c, const std::vector<double> &f, int n, std::vector<double> &out) {
        for (int i = 0; i < n - 1; ++i) out[i] = (c[i + 1] - f[i + 1]) / (c[i + 1] - c[i]);
    };
    frac(x, xu, ni, fracx);
    frac(y, yv, nj, fracy);
    frac(z, zw, nka, fracz);
    for (int j = 1; j < nj - 1; ++j)
        for (int i = 1; i < ni - 1; ++i) {
            areaij[j * ni + i] = (xu[i + 1] - xu[i]) * (yv[j + 1] - yv[j]);
            areauij[j * ni + i] = (x[i] - x[i - 1]) * (yv[j + 1] - yv[j]);
            areavij[j * ni + i] = (xu[i + 1] - xu[i]) * (y[j] - y[j - 1]);
        }
    for (int k = 1; k < nka - 1; ++k)
        for (int i = 1; i < ni - 1; ++i) {
            areaik[k * ni + i] = (xu[i + 1] - xu[i]) * (zw[k + 1] - zw[k]);
            areawik[k * ni + i] = (xu[i + 1] - xu[i]) * (z[k] - z[k - 1]);
            areauik[k * ni + i] = (x[i] - x[i - 1]) * (zw[k + 1] - zw[k]);
        }
    for (int k = 1; k < nka - 1; ++k)
        for (int j = 1; j < nj - 1; ++j) {
            areajk[k * nj + j] = (yv[j + 1] - yv[j]) * (zw[k + 1] - zw[k]);
            areavjk[k * nj + j] = (y[j] - y[j - 1]) * (zw[k + 1] - zw[k]);
            areawjk[k * nj + j] = (yv[j + 1] - yv[j]) * (z[k] - z[k - 1]);
        }
}

double HostGrid::volume(int i, int j, int k) const {
    return (xu[i + 1] - xu[i]) * (yv[j + 1] - yv[j]) * (zw[k + 1] - zw[k]);
}

// h(T) = acpa/2 T^2 + acpb T, CFDSolverManager.cpp:429-433
static double enthalpy_of(const amcfd_params &p, double t) { return 0.5 * p.acpa * t * t + p.acpb * t; }

void initial_fields(const amcfd_params &p, const HostGrid &g, InitialState &s) {
    const size_t nn = (size_t)g.ni * g.nj * g.nk;
    const double tp = p.temppreheat;
    const double hpre = enthalpy_of(p, tp);
    s.enthalpy.assign(nn, hpre);
    s.hnot.assign(nn, hpre);
    s.temp.assign(nn, tp);
    s.rho.assign(nn, p.dens);
    s.vis.assign(nn, p.viscos);
    s.diff.assign(nn, (p.thconsa * tp * tp + p.thconsb * tp + p.thconsc) / (p.acpa * tp + p.acpb));
    s.csfrac.assign(nn, 0.0);
    if (p.powderbed) {
        const double zcut = g.z[g.nk - 1] - ((double)p.numlayers * p.layerheight);
        for (int k = 0; k < g.nk; ++k) {
            double v = (g.z[k] < zcut) ? 1.0 : 0.0;
            std::fill(s.csfrac.begin() + (size_t)k * g.ni * g.nj, s.csfrac.begin() + (size_t)(k + 1) * g.ni * g.nj, v);
        }
    }
    const double hw = enthalpy_of(p, p.tempwest), he = enthalpy_of(p, p.tempeast);
    const double hn = enthalpy_of(p, p.tempnorth), hb = enthalpy_of(p, p.tempbottom);
    auto at = [&](int i, int j, int k) -> double & { return s.enthalpy[((size_t)k * g.nj + j) * g.ni + i]; };
    for (int k = 0; k < g.nk; ++k)
        for (int j = 0; j < g.nj; ++j) {
            at(0, j, k) = hw;
            at(g.ni - 1, j, k) = he;
        }
    for (int j = 0; j < g.nj; ++j)
        for (int i = 0; i < g.ni; ++i) at(i, j, 0) = hb;
    for (int k = 0; k < g.nk; ++k)
        for (int i = 0; i < g.ni; ++i) at(i, g.nj - 1, k) = hn;
}

}  // namespace amcfd
