/*
 * amcfd_oracle.c - sequential FP64 restatement of AM-CFD's per-time-step
 * SIMPLE solve.  TEST INFRASTRUCTURE ONLY (see amcfd_oracle.h).
 *
 * Arithmetic is written in the association order of the reference
 * expressions and loops run in the reference's single-thread order, so with
 * -ffp-contract=off the results match the reference to the last bit wherever
 * libm agrees.  Citations are to files under the reference checkout.
 */
#include "amcfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* std::max / std::min semantics (first argument returned on ties / NaN) */
static inline double mx(double a, double b) { return (a < b) ? b : a; }
static inline double mn(double a, double b) { return (b < a) ? b : a; }
static inline int imx(int a, int b) { return (a < b) ? b : a; }
static inline int imn(int a, int b) { return (b < a) ? b : a; }

#define SMALL 1.0e-6   /* CFDSolverManager.h:136 */
#define LARGE 1.0e20   /* CFDSolverManager.h:137 */

struct orc {
    amcfd_params p;
    double *tool, *dtt, *heatt;          /* owned copies of the tables  */
    int ni, nj, nkf;                     /* allocation extent           */
    int nk, nim1, njm1, nkm1, nim2, njm2, nkm2;
    /* geometry (CFDSolverManager.h:38-51,106-118) */
    double *x, *y, *z, *xu, *yv, *zw, *dxpwinv, *dypsinv, *dzpbinv;
    double *fracx, *fracy, *fracz;
    double *volume, *volume_u, *volume_v, *volume_w;
    double *areaij, *areauij, *areavij, *areaik, *areawik, *areauik;
    double *areajk, *areavjk, *areawjk;
    /* fields */
    double *vis, *diff, *uvel, *vvel, *wvel, *enthalpy, *temp, *pp, *hnot;
    double *fracl, *fraclnot, *sourceinput, *csfrac, *rho, *solfrac;
    double *mpphase, *mpliqphase, *unot, *vnot, *wnot, *pressure;
    double *ap, *an, *as, *ae, *aw, *at, *ab, *apnot, *dux, *dvy, *dwz, *su, *sp;
    double *heatin;
    /* species (CFDSolverManager.h:87-89) and G/R (:68-69) */
    double *conc, *concnot, *concl, *grad, *rate;
    /* energy-based free surface (FreeSurfaceManager.h): surface shift fi,
     * its previous iterate zzm, applied pressure apdis [nj*ni]; shifted
     * node heights zr and temperatures tempr [nk*nj*ni]; last bisection */
    double *fi, *zzm, *apdis, *zr, *tempr;
    double fs_xmid, fs_fmid, fs_vd;
    int fs_niter;
    /* solver scalars */
    double resoru, resorv, resorw, resorm, resorh, resorc, refmom;
    int itertot, niter, heat_it;
    /* boundary manager scalars (BoundCondManager.h:62-98) */
    double ahtoploss, heatout, accul, heatvol, ratio, tpeak;
    double fluxwest, fluxeast, fluxtop, fluxbottom, fluxnorth, fluxsouth;
    double alen, depth, width;
    int steadylength;
    int istat, jstat, kstat, iend, jend, istatp1, iendm1, jstart, istart;
    int imin, imax, jmin, jmax, kmin, kmax;
    double heatinlaser, peakhin;
    double beamposx, beamposy, beamposz, oldbeamposx, oldbeamposy, oldbeamposz;
    int powerindicator, laser_init, trackindx;
    /* driver */
    double timet, scanvel;
    int dt_state, step;
};

#define IX(s, i, j, k) ((((size_t)(k)) * (s)->nj + (size_t)(j)) * (s)->ni + (size_t)(i))

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* ------------------------------------------------------------------------ */
/* discretisation schemes, CFDSolverManager.cpp:3071-3101                    */
/* ------------------------------------------------------------------------ */
double orc_disc_power(double pe) {
    double base = 1.0 - 0.1 * pe;
    double r = base;
    for (int n = 1; n < 5; ++n) r *= base;   /* base^5 by repeated product */
    return mx(0.0, r);
}
static double disc_upwind(double pe) { (void)pe; return 1.0; }
double orc_disc_exponential(double pe) { return pe / (exp(pe) - 1.0); }

static double disc(const orc *s, double pe) {
    switch (s->p.discretization) {
    case AMCFD_DISC_POWERLAW: return orc_disc_power(pe);
    case AMCFD_DISC_EXPONENTIAL: return orc_disc_exponential(pe);
    default: return disc_upwind(pe);
    }
}

/* neighbour coefficient with convection F and diffusion D:
 * D*A(|F|/D) + max(0, +-F)  (e.g. CFDSolverManager.cpp:854-859) */
static inline double coef_pos(const orc *s, double d, double f) { return d * disc(s, fabs(f) / d) + mx(0.0, f); }
static inline double coef_neg(const orc *s, double d, double f) { return d * disc(s, fabs(f) / d) + mx(0.0, -f); }

/* ------------------------------------------------------------------------ */
/* grid, CFDSolverManager.cpp:128-286                                        */
/* ------------------------------------------------------------------------ */
static void grid_axis(int nz, const double *zone, const double *ncv, const double *powr,
                      int n, double *face, double *node) {
    face[0] = 0.0;
    face[1] = 0.0;
    int ist = 2;
    double statloc = 0.0;
    for (int zi = 0; zi < nz; ++zi) {
        for (int c = 0; c < ncv[zi]; ++c) {
            double frac = (double)(c + 1) / ncv[zi];
            double term = (powr[zi] >= 0.0) ? pow(frac, powr[zi])
                                            : 1.0 - pow(1.0 - frac, -powr[zi]);
            face[c + ist] = statloc + zone[zi] * term;
        }
        ist += (int)ncv[zi];
        statloc += zone[zi];
    }
    for (int i = 0; i < n - 1; ++i) node[i] = (face[i + 1] + face[i]) * 0.5;
    node[n - 1] = face[n - 1];
}

/* The z extent is the active nk: in powder mode the reference regenerates the
 * grid (moving mesh) with the current nk_, leaving entries above it as they
 * were; at set-up nk is the full extent. */
static void generate_grid(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj, nk = s->nk;
    grid_axis(p->nzx, p->xzone, p->ncvx, p->powrx, ni, s->xu, s->x);
    grid_axis(p->nzy, p->yzone, p->ncvy, p->powry, nj, s->yv, s->y);
    grid_axis(p->nzz, p->zzone, p->ncvz, p->powrz, nk, s->zw, s->z);
    for (int i = 1; i < ni; ++i) s->dxpwinv[i] = 1.0 / (s->x[i] - s->x[i - 1]);
    for (int j = 1; j < nj; ++j) s->dypsinv[j] = 1.0 / (s->y[j] - s->y[j - 1]);
    for (int k = 1; k < nk; ++k) s->dzpbinv[k] = 1.0 / (s->z[k] - s->z[k - 1]);
    for (int i = 0; i < ni - 1; ++i) s->fracx[i] = (s->x[i + 1] - s->xu[i + 1]) / (s->x[i + 1] - s->x[i]);
    for (int j = 0; j < nj - 1; ++j) s->fracy[j] = (s->y[j + 1] - s->yv[j + 1]) / (s->y[j + 1] - s->y[j]);
    for (int k = 0; k < nk - 1; ++k) s->fracz[k] = (s->z[k + 1] - s->zw[k + 1]) / (s->z[k + 1] - s->z[k]);
    for (int k = 1; k < nk - 1; ++k) {
        double dzc = s->zw[k + 1] - s->zw[k], dzs = s->z[k] - s->z[k - 1];
        for (int j = 1; j < nj - 1; ++j) {
            double dyc = s->yv[j + 1] - s->yv[j], dys = s->y[j] - s->y[j - 1];
            for (int i = 1; i < ni - 1; ++i) {
                double dxc = s->xu[i + 1] - s->xu[i], dxs = s->x[i] - s->x[i - 1];
                size_t c = ((size_t)k * nj + j) * ni + i;
                s->volume[c] = dxc * dyc * dzc;
                s->volume_u[c] = dxs * dyc * dzc;
                s->volume_v[c] = dxc * dys * dzc;
                s->volume_w[c] = dxc * dyc * dzs;
            }
        }
    }
    for (int j = 1; j < nj - 1; ++j)
        for (int i = 1; i < ni - 1; ++i) {
            s->areaij[j * ni + i] = (s->xu[i + 1] - s->xu[i]) * (s->yv[j + 1] - s->yv[j]);
            s->areauij[j * ni + i] = (s->x[i] - s->x[i - 1]) * (s->yv[j + 1] - s->yv[j]);
            s->areavij[j * ni + i] = (s->xu[i + 1] - s->xu[i]) * (s->y[j] - s->y[j - 1]);
        }
    for (int k = 1; k < nk - 1; ++k)
        for (int i = 1; i < ni - 1; ++i) {
            s->areaik[k * ni + i] = (s->xu[i + 1] - s->xu[i]) * (s->zw[k + 1] - s->zw[k]);
            s->areawik[k * ni + i] = (s->xu[i + 1] - s->xu[i]) * (s->z[k] - s->z[k - 1]);
            s->areauik[k * ni + i] = (s->x[i] - s->x[i - 1]) * (s->zw[k + 1] - s->zw[k]);
        }
    for (int k = 1; k < nk - 1; ++k)
        for (int j = 1; j < nj - 1; ++j) {
            s->areajk[k * nj + j] = (s->yv[j + 1] - s->yv[j]) * (s->zw[k + 1] - s->zw[k]);
            s->areavjk[k * nj + j] = (s->y[j] - s->y[j - 1]) * (s->zw[k + 1] - s->zw[k]);
            s->areawjk[k * nj + j] = (s->yv[j + 1] - s->yv[j]) * (s->z[k] - s->z[k - 1]);
        }
}

static void set_active_nk(orc *s, int nk) {
    s->nk = nk;
    s->nkm1 = nk - 1;
    s->nkm2 = nk - 2;
}

/* linear interpolation, CFDSolverManager.cpp:3057-3066 */
static double lerp_ref(double x0, double x1, double y0, double y1, double x) {
    double dy = y1 - y0, dx = x1 - x0;
    return y0 + (x - x0) * (dy / dx);
}

/* Moving mesh, CFDSolverManager.cpp:60-124 (zone shift) + 345-346 and
 * 2809-3051 (remap).  The zones either side of the moving one take up the
 * beam's displacement since the last step; the grid is rebuilt; then twelve
 * fields (fourteen with species) are re-interpolated in place, first along x (node i from i and
 * i+dir), then along y (j and j+dir), wherever a node moved by more than
 * 1e-6.  Sweeping i (j) upwards in place makes a backward move (dir = -1) a
 * recurrence on the already-remapped neighbour. */
static void move_mesh(orc *s) {
    amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj, nk = s->nk;
    double *oldx = (double *)malloc(sizeof(double) * ni), *oldy = (double *)malloc(sizeof(double) * nj);
    memcpy(oldx, s->x, sizeof(double) * ni);
    memcpy(oldy, s->y, sizeof(double) * nj);
    double dx = s->beamposx - s->oldbeamposx, dy = s->beamposy - s->oldbeamposy;
    int dirx = (dx < 0) ? -1 : 1, diry = (dy < 0) ? -1 : 1;
    int ix = p->xmoveindex, iy = p->ymoveindex;
    p->xzone[ix - 1] += dx;
    p->xzone[ix + 1] -= dx;
    /* until the pool length is steady the y zones are picked by the x index */
    int iyz = s->steadylength ? iy : ix;
    p->yzone[iyz - 1] += dy;
    p->yzone[iyz + 1] -= dy;
    generate_grid(s);
    double *f[14] = {s->enthalpy, s->hnot, s->temp, s->fracl, s->uvel, s->vvel,
                     s->wvel, s->unot, s->vnot, s->wnot, s->pressure, s->pp, s->conc, s->concnot};
    const int nf = p->species ? 14 : 12;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i) {
                if (!(fabs(oldx[i] - s->x[i]) > 1e-6)) continue;
                size_t c = IX(s, i, j, k);
                for (int a = 0; a < nf; ++a)
                    f[a][c] = lerp_ref(oldx[i], oldx[i + dirx], f[a][c], f[a][c + dirx], s->x[i]);
            }
    const long sy = (long)diry * ni;
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj - 1; ++j)
            for (int i = 1; i < ni - 1; ++i) {
                if (!(fabs(oldy[j] - s->y[j]) > 1e-6)) continue;
                size_t c = IX(s, i, j, k);
                for (int a = 0; a < nf; ++a)
                    f[a][c] = lerp_ref(oldy[j], oldy[j + diry], f[a][c], f[a][(size_t)((long)c + sy)], s->y[j]);
            }
    free(oldx);
    free(oldy);
}

/* ------------------------------------------------------------------------ */
/* initial state, CFDSolverManager.cpp:402-590                               */
/* ------------------------------------------------------------------------ */
static double h_of_t(const amcfd_params *p, double t) { return 0.5 * p->acpa * t * t + p->acpb * t; }

static void initialize(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj, nk = s->nkf;
    double tp = p->temppreheat;
    double hpre = h_of_t(p, tp), hwest = h_of_t(p, p->tempwest), heast = h_of_t(p, p->tempeast);
    double hnorth = h_of_t(p, p->tempnorth), hbottom = h_of_t(p, p->tempbottom);
    double d0 = (p->thconsa * tp * tp + p->thconsb * tp + p->thconsc) / (p->acpa * tp + p->acpb);
    size_t nn = (size_t)ni * nj * nk;
    double ztop = s->z[nk - 1] - ((double)p->numlayers * p->layerheight);
    for (size_t c = 0; c < nn; ++c) {
        s->vis[c] = p->viscos;
        s->rho[c] = p->dens;
        s->diff[c] = d0;
        s->enthalpy[c] = hpre;
        s->hnot[c] = hpre;
        s->temp[c] = tp;
        if (p->species) {   /* CFDSolverManager.cpp:474-479 */
            s->conc[c] = p->initconcentration;
            s->concnot[c] = p->initconcentration;
            s->concl[c] = p->initconcentration;
        }
        if (p->powderbed) {
            int k = (int)(c / ((size_t)ni * nj));
            s->csfrac[c] = (s->z[k] < ztop) ? 1.0 : 0.0;
        }
    }
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j) {
            s->enthalpy[IX(s, 0, j, k)] = hwest;
            s->enthalpy[IX(s, ni - 1, j, k)] = heast;
        }
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) s->enthalpy[IX(s, i, j, 0)] = hbottom;
    for (int k = 0; k < nk; ++k)
        for (int i = 0; i < ni; ++i) s->enthalpy[IX(s, i, nj - 1, k)] = hnorth;
}

/* ------------------------------------------------------------------------ */
/* laser, BoundCondManager.cpp:68-275                                        */
/* ------------------------------------------------------------------------ */
static void laser_coords(orc *s) {
    const amcfd_params *p = &s->p;
    if (!p->udtoolpath) {
        s->oldbeamposx = s->beamposx;
        if (s->timet <= p->lasertime) {
            s->beamposx = p->xstart + s->timet * s->scanvel;
            if (s->laser_init) {
                s->oldbeamposx = s->beamposx;
                s->laser_init = 0;
            }
        } else {
            s->beamposx = p->xstart + p->lasertime * s->scanvel;
            s->powerindicator = 0;
        }
        s->beamposy = p->ystart;
        return;
    }
    /* toolpath: piecewise-linear position between rows, BoundCondManager.cpp:203-273 */
    const double small = 1.0e-8;
    double tx = s->beamposx, ty = s->beamposy, tz = s->beamposz, dty = 0.0;
    s->oldbeamposx = s->beamposx;
    s->oldbeamposy = s->beamposy;
    s->oldbeamposz = s->beamposz;
    int oldpi = s->powerindicator;
    for (int r = 1; r < p->n_toolpath; ++r) {
        const double *b = s->tool + 6 * r, *a = s->tool + 6 * (r - 1);
        if (s->timet <= b[0] + small) {
            double ratio = (s->timet - a[0]) / (b[0] - a[0]);
            double dtx = b[1] - a[1];
            dty = b[2] - a[2];
            double dtz = b[3] - a[3];
            tx = ratio * dtx + a[1];
            ty = ratio * dty + a[2];
            tz = ratio * dtz + a[3];
            s->powerindicator = (int)b[4];
            if (s->laser_init) {
                s->oldbeamposx = a[1];
                s->oldbeamposy = a[2];
                s->oldbeamposz = a[3];
                oldpi = (int)a[4];
                s->laser_init = 0;
            }
            break;
        }
    }
    s->beamposx = tx;
    s->beamposy = ty;
    s->beamposz = tz;
    double dx = s->beamposx - s->oldbeamposx, dy = s->beamposy - s->oldbeamposy,
           dz = s->beamposz - s->oldbeamposz;
    s->scanvel = sqrt(dx * dx + dy * dy + dz * dz) / p->delt;
    if (oldpi - s->powerindicator > 0 && dty > small) s->trackindx++;
    if (s->powerindicator - oldpi > 0 && p->powderbed) {
        s->trackindx++;
        /* more layers than *POWDER_BED_MODE declares would run past the grid */
        if (s->nk + p->ncvzlayer <= s->nkf) set_active_nk(s, s->nk + p->ncvzlayer);
    }
}

static int nearest_index(const double *c, int nlast, double pos) {
    int out = 0;
    for (int i = 1; i < nlast; ++i) {
        if (pos <= c[i]) break;
        out = i;
    }
    if (fabs(pos - c[out + 1]) < fabs(pos - c[out])) out++;
    return out;
}

static void update_laser(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj;
    if (p->steady) s->beamposx = p->xstart;
    else laser_coords(s);
    s->istart = nearest_index(s->x, s->nim1, s->beamposx);
    s->jstart = nearest_index(s->y, s->njm1, s->beamposy);
    s->heatinlaser = 0.0;
    double rb2 = p->alasrb * p->alasrb;
    double varlas = p->alaspow * p->alaseta;
    s->peakhin = p->alasfact * varlas / (M_PI * rb2);
    for (int j = 0; j < nj; ++j) {
        double yd = s->beamposy - s->y[j];
        for (int i = 0; i < ni; ++i) {
            double xd = s->beamposx - s->x[i];
            double d2 = xd * xd + yd * yd;
            double q = s->peakhin * exp(-p->alasfact * d2 / rb2) * (double)s->powerindicator;
            if (p->alasfact < 1.0e-6 && d2 < rb2) q = varlas / (M_PI * rb2) * (double)s->powerindicator;
            s->heatin[j * ni + i] = q;
            s->heatinlaser += s->areaij[j * ni + i] * q;
        }
    }
    s->imin = s->istart - 2;
    s->imax = s->istart + 2;
    s->jmin = s->jstart - 2;
    s->jmax = s->jstart + 2;
    s->kmin = s->nkm2 - 3;
    s->kmax = s->nkm2;
}

/* ------------------------------------------------------------------------ */
/* boundary conditions, BoundCondManager.cpp:452-701                         */
/* ------------------------------------------------------------------------ */
static double surface_tension_gradient(const orc *s, double t, double a0) {
    /* Fe-O surfactant model, BoundCondManager.cpp:706-741 */
    const double A = 4.3e-4, R = 8314, gamma0 = 1.943, Ts = 2.03e-8, k1 = 1.38e-2, H0 = -1.46e8;
    (void)gamma0;
    double K = k1 * exp(-H0 / R / t);
    double g = (-A - R * Ts * log(1 + K * a0) - (K * a0 / (1 + K * a0)) * (Ts * H0 / t)) * 1;
    if (t >= 2500) g = 0;
    if (a0 < 0) g = s->p.dgdtp;
    return g;
}

static double pow4(double b) { return b * b * b * b; }

static void bc_enthalpy(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nim1 = s->nim1, nim2 = s->nim2, njm1 = s->njm1, njm2 = s->njm2;
    const int nkm1 = s->nkm1, nkm2 = s->nkm2;
    double hevap = 0.0;
    s->ahtoploss = 0.0;
    for (int j = 1; j < njm1; ++j)
        for (int i = 1; i < nim1; ++i) {
            double t = s->temp[IX(s, i, j, nkm1)];
            if (p->isevap) {
                double psat = p->pressamb * exp(-p->evaplatent * p->molarmass / p->gasconst *
                                                (1.0 / t - 1.0 / p->evaptemp));
                double mdot = p->evapfact * psat * sqrt(p->molarmass / (2.0 * M_PI * p->gasconst * t));
                hevap = mdot * p->evaplatent;
            }
            double hrad = p->emiss * p->sigm * (pow4(t) - pow4(p->tempamb));
            double c1 = s->diff[IX(s, i, j, nkm2)] * s->dzpbinv[nkm1];
            s->enthalpy[IX(s, i, j, nkm1)] =
                s->enthalpy[IX(s, i, j, nkm2)] + (s->heatin[j * ni + i] - hrad - hrad - hevap) / c1;
            s->ahtoploss += (hrad + hrad + hevap) * s->areaij[j * ni + i];
        }
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < njm1; ++j) {
            double q = p->htci * (s->temp[IX(s, 0, j, k)] - p->tempamb);
            double c1 = s->diff[IX(s, 1, j, k)] * s->dxpwinv[1];
            s->enthalpy[IX(s, 0, j, k)] = s->enthalpy[IX(s, 1, j, k)] - q / c1;
            q = p->htci * (s->temp[IX(s, nim1, j, k)] - p->tempamb);
            c1 = s->diff[IX(s, nim2, j, k)] * s->dxpwinv[nim2];
            s->enthalpy[IX(s, nim1, j, k)] = s->enthalpy[IX(s, nim2, j, k)] - q / c1;
        }
    for (int k = 1; k < nkm1; ++k)
        for (int i = 1; i < nim1; ++i) {
            double q = p->htcj * (s->temp[IX(s, i, 0, k)] - p->tempamb);
            double c1 = s->diff[IX(s, i, 1, k)] * s->dypsinv[1];
            s->enthalpy[IX(s, i, 0, k)] = s->enthalpy[IX(s, i, 1, k)] - q / c1;
            q = p->htcj * (s->temp[IX(s, i, njm1, k)] - p->tempamb);
            c1 = s->diff[IX(s, i, njm2, k)] * s->dypsinv[njm2];
            s->enthalpy[IX(s, i, njm1, k)] = s->enthalpy[IX(s, i, njm2, k)] - q / c1;
        }
}

/* Marangoni top ghost + box side zeroing for u (ivar 1) or v (ivar 2),
 * BoundCondManager.cpp:472-536 */
static void bc_tangential(orc *s, int ivar) {
    const int ni = s->ni, nkm1 = s->nkm1, nkm2 = s->nkm2;
    double *phi = (ivar == 1) ? s->uvel : s->vvel;
    for (int j = s->jstat; j <= s->jend; ++j)
        for (int i = s->istatp1; i <= s->iendm1; ++i) {
            size_t top = IX(s, i, j, nkm1);
            double grad, flf, visf;
            if (ivar == 1) {
                size_t w = IX(s, i - 1, j, nkm1);
                grad = (s->temp[top] - s->temp[w]) * s->dxpwinv[i];
                flf = s->fracl[top] * (1.0 - s->fracx[i - 1]) + s->fracl[w] * s->fracx[i - 1];
                double vp = s->vis[IX(s, i, j, nkm2)], vw = s->vis[IX(s, i - 1, j, nkm2)];
                visf = vp * vw / (vp * (1.0 - s->fracx[i - 1]) + vw * s->fracx[i - 1]);
            } else {
                size_t so = IX(s, i, j - 1, nkm1);
                grad = (s->temp[top] - s->temp[so]) * s->dypsinv[j];
                flf = s->fracl[top] * (1.0 - s->fracy[j - 1]) + s->fracl[so] * s->fracy[j - 1];
                double vp = s->vis[IX(s, i, j, nkm2)], vs = s->vis[IX(s, i, j - 1, nkm2)];
                visf = vp * vs / (vp * (1.0 - s->fracy[j - 1]) + vs * s->fracy[j - 1]);
            }
            double term = flf * surface_tension_gradient(s, s->temp[top], s->p.a_oxygen) * grad /
                          (visf * s->dzpbinv[nkm1]);
            phi[top] = phi[IX(s, i, j, nkm2)] + term;
        }
    (void)ni;
}

static void bc_box_sides(orc *s, double *phi) {
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j) {
            phi[IX(s, s->istat, j, k)] = 0.0;
            phi[IX(s, s->iend, j, k)] = 0.0;
        }
}

/* species boundary values, BoundCondManager.cpp:655-699: top from the
 * surface flux where the surface is liquid, fixed bottom, then zero-gradient
 * north/south and east/west over whole planes (corners take the last copy) */
static void bc_species(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj, nk = s->nk;
    double *c = s->conc;
    const double dl = p->dens * p->massdiff;
    for (int j = 1; j < s->njm1; ++j)
        for (int i = 1; i < s->nim1; ++i) {
            size_t t = IX(s, i, j, s->nkm1), b = IX(s, i, j, s->nkm2);
            if (s->temp[t] > p->tliquid) c[t] = 2.0 * (13.0 - c[b]) / s->dzpbinv[s->nkm1] / dl + c[b];
            else c[t] = c[b];
            c[IX(s, i, j, 0)] = p->initconcentration;
        }
    for (int k = 0; k < nk; ++k)
        for (int i = 0; i < ni; ++i) {
            c[IX(s, i, 0, k)] = c[IX(s, i, 1, k)];
            c[IX(s, i, s->njm1, k)] = c[IX(s, i, s->njm1 - 1, k)];
        }
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j) {
            c[IX(s, 0, j, k)] = c[IX(s, 1, j, k)];
            c[IX(s, s->nim1, j, k)] = c[IX(s, s->nim1 - 1, j, k)];
        }
}

static void apply_bcs(orc *s, int ivar) {
    switch (ivar) {
    case 1: bc_tangential(s, 1); bc_box_sides(s, s->uvel); break;
    case 2: bc_tangential(s, 2); bc_box_sides(s, s->vvel); break;
    case 3: bc_box_sides(s, s->wvel); break;
    case 4: bc_box_sides(s, s->pp); break;
    case 5: bc_enthalpy(s); break;
    }
}

/* ------------------------------------------------------------------------ */
/* energy coefficients, CFDSolverManager.cpp:1157-1263                       */
/* ------------------------------------------------------------------------ */
static void discretize_energy(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const int nim1 = s->nim1, njm1 = s->njm1, nkm1 = s->nkm1, nim2 = s->nim2, njm2 = s->njm2,
              nkm2 = s->nkm2;
    const double delt = s->p.delt;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < njm1; ++j)
            for (int i = 1; i < nim1; ++i) {
                size_t c = IX(s, i, j, k);
                double vn = s->vvel[c + sj], ue = s->uvel[c + 1], wt = s->wvel[c + sk];
                double densbydt = s->rho[c] / delt;
                double fn = s->rho[c + sj] * vn * s->areaik[k * ni + i];
                double fe = s->rho[c + 1] * ue * s->areajk[k * nj + j];
                double ft = s->rho[c + sk] * wt * s->areaij[j * ni + i];
                double dP = s->diff[c];
                double difn = dP * s->diff[c + sj] / ((1.0 - s->fracy[j]) * s->diff[c + sj] + s->fracy[j] * dP);
                double dife = dP * s->diff[c + 1] / ((1.0 - s->fracx[i]) * s->diff[c + 1] + s->fracx[i] * dP);
                double dift = dP * s->diff[c + sk] / ((1.0 - s->fracz[k]) * s->diff[c + sk] + s->fracz[k] * dP);
                double dn = difn * s->areaik[k * ni + i] * s->dypsinv[j + 1];
                double de = dife * s->areajk[k * nj + j] * s->dxpwinv[i + 1];
                double dt = dift * s->areaij[j * ni + i] * s->dzpbinv[k + 1];
                s->an[c] = coef_neg(s, dn, fn);
                s->as[c + sj] = coef_pos(s, dn, fn);
                s->ae[c] = coef_neg(s, de, fe);
                s->aw[c + 1] = coef_pos(s, de, fe);
                s->at[c] = coef_neg(s, dt, ft);
                s->ab[c + sk] = coef_pos(s, dt, ft);
                s->apnot[c] = densbydt * s->volume[c];
                s->sp[c] = 0.0;
                s->su[c] = s->apnot[c] * s->hnot[c];
            }
    /* boundary faces use the boundary-adjacent diffusivity only */
    for (int k = 1; k < nkm1; ++k)
        for (int i = 1; i < nim1; ++i) {
            size_t c1 = IX(s, i, 1, k), cn = IX(s, i, njm2, k), cb = IX(s, i, njm1, k);
            double fs = s->rho[c1] * s->vvel[c1] * s->areaik[k * ni + i];
            double ds = s->diff[c1] * s->areaik[k * ni + i] * s->dypsinv[1];
            s->as[c1] = coef_pos(s, ds, fs);
            double fn = s->rho[cb] * s->vvel[cb] * s->areaik[k * ni + i];
            double dn = s->diff[cn] * s->areaik[k * ni + i] * s->dypsinv[njm1];
            s->an[cn] = coef_neg(s, dn, fn);
        }
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < njm1; ++j) {
            size_t c1 = IX(s, 1, j, k), cn = IX(s, nim2, j, k), cb = IX(s, nim1, j, k);
            double fw = s->rho[c1] * s->uvel[c1] * s->areajk[k * nj + j];
            double dw = s->diff[c1] * s->areajk[k * nj + j] * s->dxpwinv[1];
            s->aw[c1] = coef_pos(s, dw, fw);
            double fe = s->rho[cb] * s->uvel[cb] * s->areajk[k * nj + j];
            double de = s->diff[cn] * s->areajk[k * nj + j] * s->dxpwinv[nim1];
            s->ae[cn] = coef_neg(s, de, fe);
        }
    for (int j = 1; j < njm1; ++j)
        for (int i = 1; i < nim1; ++i) {
            size_t c1 = IX(s, i, j, 1), cn = IX(s, i, j, nkm2), cb = IX(s, i, j, nkm1);
            double fb = s->rho[c1] * s->wvel[c1] * s->areaij[j * ni + i];
            double db = s->diff[c1] * s->areaij[j * ni + i] * s->dzpbinv[1];
            s->ab[c1] = coef_pos(s, db, fb);
            double ft = s->rho[cb] * s->wvel[cb] * s->areaij[j * ni + i];
            double dt = s->diff[cn] * s->areaij[j * ni + i] * s->dzpbinv[nkm1];
            s->at[cn] = coef_neg(s, dt, ft);
        }
}

/* heat-source table lookup + energy sources, CFDSolverManager.cpp:1677-1843 */
static void source_energy(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj;
    const int nim1 = s->nim1, njm1 = s->njm1, nkm1 = s->nkm1, nim2 = s->nim2, njm2 = s->njm2,
              nkm2 = s->nkm2;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    double avolfact = 0, avolpow = 0, apowseta = 0, heatthick = 0, heatrb = 0;
    if (p->volheatsource && !p->volheatsourcetable) {
        avolfact = p->avolfact;
        avolpow = p->avolpow;
        apowseta = p->apowseta;
        heatthick = p->heatthick;
        heatrb = p->heatrb;
    } else if (p->n_heat_table > 0) {
        if (s->timet > s->heatt[6 * s->heat_it] && s->heat_it < p->n_heat_table - 1) s->heat_it += 1;
        const double *r = s->heatt + 6 * s->heat_it;
        avolpow = r[1];
        apowseta = r[2];
        heatthick = r[3];
        heatrb = r[4];
        avolfact = r[5];
    }
    const double hlatnt = p->hlatnt;
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < njm1; ++j)
            for (int i = 1; i < nim1; ++i) {
                size_t c = IX(s, i, j, k);
                if (s->z[nkm1] - s->z[k] <= heatthick) {
                    double bx = s->beamposx - s->x[i], by = s->beamposy - s->y[j];
                    s->sourceinput[c] = avolpow * avolfact / M_PI / (heatrb * heatrb) / heatthick * apowseta *
                                        exp(-avolfact / (heatrb * heatrb) * ((bx * bx) + (by * by))) *
                                        (double)s->powerindicator;
                    s->su[c] += s->volume[c] * s->sourceinput[c];
                } else {
                    s->sourceinput[c] = 0.0;
                }
                double v1 = hlatnt * s->rho[c] / p->delt;
                double volht = s->volume[c] * v1;
                s->su[c] -= volht * (s->fracl[c] - s->fraclnot[c]);
                const double *fl = s->fracl;
                double ue = s->uvel[c + 1], uw = s->uvel[c];
                double vn = s->vvel[c + sj], vs = s->vvel[c];
                double wt = s->wvel[c + sk], wb = s->wvel[c];
                double flew = s->areajk[k * nj + j] * (mx(uw, 0.0) * fl[c - 1] - mx(-uw, 0.0) * fl[c] +
                                                       mx(-ue, 0.0) * fl[c + 1] - mx(ue, 0.0) * fl[c]);
                double flns = s->areaik[k * ni + i] * (mx(vs, 0.0) * fl[c - sj] - mx(-vs, 0.0) * fl[c] +
                                                       mx(-vn, 0.0) * fl[c + sj] - mx(vn, 0.0) * fl[c]);
                double fltb = s->areaij[j * ni + i] * (mx(wb, 0.0) * fl[c - sk] - mx(-wb, 0.0) * fl[c] +
                                                       mx(-wt, 0.0) * fl[c + sk] - mx(wt, 0.0) * fl[c]);
                s->su[c] += s->rho[c] * hlatnt * (flew + flns + fltb);
            }
    /* fold fixed boundary values into the source: bottom/top, south/north, west/east */
    for (int j = 1; j < njm1; ++j)
        for (int i = 1; i < nim1; ++i) {
            size_t c = IX(s, i, j, 1);
            s->su[c] += s->ab[c] * s->enthalpy[c - sk];
            s->sp[c] -= s->ab[c];
            s->ab[c] = 0.0;
            c = IX(s, i, j, nkm2);
            s->su[c] += s->at[c] * s->enthalpy[c + sk];
            s->sp[c] -= s->at[c];
            s->at[c] = 0.0;
        }
    for (int k = 1; k < nkm1; ++k)
        for (int i = 1; i < nim1; ++i) {
            size_t c = IX(s, i, 1, k);
            s->su[c] += s->as[c] * s->enthalpy[c - sj];
            s->sp[c] -= s->as[c];
            s->as[c] = 0.0;
            c = IX(s, i, njm2, k);
            s->su[c] += s->an[c] * s->enthalpy[c + sj];
            s->sp[c] -= s->an[c];
            s->an[c] = 0.0;
        }
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < njm1; ++j) {
            size_t c = IX(s, 1, j, k);
            s->su[c] += s->aw[c] * s->enthalpy[c - 1];
            s->sp[c] -= s->aw[c];
            s->aw[c] = 0.0;
            c = IX(s, nim2, j, k);
            s->su[c] += s->ae[c] * s->enthalpy[c + 1];
            s->sp[c] -= s->ae[c];
            s->ae[c] = 0.0;
        }
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < njm1; ++j)
            for (int i = 1; i < nim1; ++i) {
                size_t c = IX(s, i, j, k);
                s->ap[c] = s->an[c] + s->as[c] + s->ae[c] + s->aw[c] + s->at[c] + s->ab[c] + s->apnot[c] - s->sp[c];
                s->ap[c] /= p->urfh;
                s->su[c] += (1.0 - p->urfh) * s->ap[c] * s->enthalpy[c];
            }
}

/* species coefficients, CFDSolverManager.cpp:1265-1367: the cell's own
 * diffusivity on every face (no harmonic mean; 1e-11 in solid), convection
 * with twice the face velocity; each cell writes its north/east/top
 * coefficient and its neighbour's south/west/bottom one, the first row,
 * column and plane get theirs from a separate loop */
static void discretize_species(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj, n = (size_t)ni * nj * s->nk;
    double *ds = (double *)malloc(sizeof(double) * (n ? n : 1));
    for (size_t c = 0; c < n; ++c) ds[c] = (s->temp[c] <= p->tsolid) ? 1e-11 : p->massdiff * s->rho[c];
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j)
            for (int i = 1; i < s->nim1; ++i) {
                size_t c = IX(s, i, j, k);
                double aik = s->areaik[k * ni + i], ajk = s->areajk[k * nj + j], aij = s->areaij[j * ni + i];
                double vn = s->vvel[c + sj] * 2.0, ue = s->uvel[c + 1] * 2.0, wt = s->wvel[c + sk] * 2.0;
                double fn = s->rho[c + sj] * vn * aik, fe = s->rho[c + 1] * ue * ajk, ft = s->rho[c + sk] * wt * aij;
                double dn = ds[c] * aik * s->dypsinv[j + 1], de = ds[c] * ajk * s->dxpwinv[i + 1];
                double dt = ds[c] * aij * s->dzpbinv[k + 1];
                double densbydt = s->rho[c] / p->delt;
                s->an[c] = coef_neg(s, dn, fn);
                s->as[c + sj] = coef_pos(s, dn, fn);
                s->ae[c] = coef_neg(s, de, fe);
                s->aw[c + 1] = coef_pos(s, de, fe);
                s->at[c] = coef_neg(s, dt, ft);
                s->ab[c + sk] = coef_pos(s, dt, ft);
                s->apnot[c] = densbydt * s->volume[c];
                s->sp[c] = 0.0;
                s->su[c] = s->apnot[c] * s->concnot[c];
            }
    for (int k = 1; k < s->nkm1; ++k)
        for (int i = 1; i < s->nim1; ++i) {
            size_t c = IX(s, i, 1, k);
            double aik = s->areaik[k * ni + i];
            double vs = s->vvel[c] * 2.0, fs = s->rho[c] * vs * aik, dd = ds[c] * aik * s->dypsinv[1];
            s->as[c] = coef_pos(s, dd, fs);
        }
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j) {
            size_t c = IX(s, 1, j, k);
            double ajk = s->areajk[k * nj + j];
            double uw = s->uvel[c] * 2.0, fw = s->rho[c] * uw * ajk, dd = ds[c] * ajk * s->dxpwinv[1];
            s->aw[c] = coef_pos(s, dd, fw);
        }
    for (int j = 1; j < s->njm1; ++j)
        for (int i = 1; i < s->nim1; ++i) {
            size_t c = IX(s, i, j, 1);
            double aij = s->areaij[j * ni + i];
            double wb = s->wvel[c] * 2.0, fb = s->rho[c] * wb * aij, dd = ds[c] * aij * s->dzpbinv[1];
            s->ab[c] = coef_pos(s, dd, fb);
        }
    free(ds);
}

/* boundary folding, ap and relaxation, CFDSolverManager.cpp:1844-1924 */
static void source_species(orc *s) {
    const amcfd_params *p = &s->p;
    const double *c = s->conc;
    for (int j = 1; j < s->njm1; ++j)
        for (int i = 1; i < s->nim1; ++i) {
            size_t b = IX(s, i, j, 1), t = IX(s, i, j, s->nkm2);
            s->su[b] += s->ab[b] * c[IX(s, i, j, 0)];
            s->sp[b] -= s->ab[b];
            s->ab[b] = 0.0;
            s->su[t] += s->at[t] * c[IX(s, i, j, s->nkm1)];
            s->sp[t] -= s->at[t];
            s->at[t] = 0.0;
        }
    for (int k = 1; k < s->nkm1; ++k)
        for (int i = 1; i < s->nim1; ++i) {
            size_t a = IX(s, i, 1, k), b = IX(s, i, s->njm2, k);
            s->su[a] += s->as[a] * c[IX(s, i, 0, k)];
            s->sp[a] -= s->as[a];
            s->as[a] = 0.0;
            s->su[b] += s->an[b] * c[IX(s, i, s->njm1, k)];
            s->sp[b] -= s->an[b];
            s->an[b] = 0.0;
        }
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j) {
            size_t a = IX(s, 1, j, k), b = IX(s, s->nim2, j, k);
            s->su[a] += s->aw[a] * c[IX(s, 0, j, k)];
            s->sp[a] -= s->aw[a];
            s->aw[a] = 0.0;
            s->su[b] += s->ae[b] * c[IX(s, s->nim1, j, k)];
            s->sp[b] -= s->ae[b];
            s->ae[b] = 0.0;
        }
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j)
            for (int i = 1; i < s->nim1; ++i) {
                size_t q = IX(s, i, j, k);
                double ap = s->an[q] + s->as[q] + s->ae[q] + s->aw[q] + s->at[q] + s->ab[q] + s->apnot[q] - s->sp[q];
                ap /= p->urfc;
                s->ap[q] = ap;
                s->su[q] += (1.0 - p->urfc) * ap * c[q];
            }
}

/* L1 species residual, CFDSolverManager.cpp:2102-2130 */
static void residual_species(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    double sumc = 0.0, sumd = 0.0;
    const double *h = s->conc;
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j)
            for (int i = 1; i < s->nim1; ++i) {
                size_t c = IX(s, i, j, k);
                double r = (s->an[c] * h[c + sj] + s->as[c] * h[c - sj] + s->ae[c] * h[c + 1] + s->aw[c] * h[c - 1] +
                            s->at[c] * h[c + sk] + s->ab[c] * h[c - sk] + s->su[c]) / s->ap[c] - h[c];
                sumd += fabs(r);
                sumc += fabs(h[c]);
            }
    s->resorc = sumd / (sumc + SMALL);
}

/* L1 enthalpy residual, CFDSolverManager.cpp:2073-2101 */
static void residual_energy(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    double sumh = 0.0, sumd = 0.0;
    const double *h = s->enthalpy;
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j)
            for (int i = 1; i < s->nim1; ++i) {
                size_t c = IX(s, i, j, k);
                double r = (s->an[c] * h[c + sj] + s->as[c] * h[c - sj] + s->ae[c] * h[c + 1] + s->aw[c] * h[c - 1] +
                            s->at[c] * h[c + sk] + s->ab[c] * h[c - sk] + s->su[c]) / s->ap[c] - h[c];
                sumd += fabs(r);
                sumh += fabs(h[c]);
            }
    s->resorh = sumd / (sumh + SMALL);
}

/* phase-wise properties, CFDSolverManager.cpp:2137-2265 */
static void material_props(orc *s) {
    const amcfd_params *p = &s->p;
    size_t nn = (size_t)s->ni * s->nj * s->nk;
    for (size_t c = 0; c < nn; ++c) {
        double t = s->temp[c];
        double diffl = (p->thconla * t + p->thconlb) / p->acpl;
        s->diff[c] = diffl;
        s->rho[c] = p->denl;
        s->vis[c] = p->viscos;
        if (t >= p->tliquid) continue;
        double diffs, rhos;
        if (!p->powderbed) {
            diffs = (p->thconsa * (t * t) + p->thconsb * t + p->thconsc) / (p->acpa * t + p->acpb);
            rhos = p->dens;
        } else {
            double cps = p->acpa * t + p->acpb;
            double conds = p->thconsa * (t * t) + p->thconsb * t + p->thconsc;
            double diffp = (p->powderthcona * t + p->powderthconb) / (p->powdercpa * t + p->powdercpb);
            diffs = (1.0 - s->csfrac[c]) * diffp + s->csfrac[c] * conds / cps;
            rhos = (1.0 - s->csfrac[c]) * p->powderden + s->csfrac[c] * p->dens;
        }
        s->diff[c] = diffs;
        s->rho[c] = rhos;
        s->vis[c] = 1e10;
        if (t <= p->tsolid) continue;
        s->diff[c] = s->fracl[c] * diffl + (1.0 - s->fracl[c]) * diffs;
        s->rho[c] = s->fracl[c] * p->denl + (1.0 - s->fracl[c]) * rhos;
        s->vis[c] = p->viscos;
    }
}

/* x-direction block correction, CFDSolverManager.cpp:2270-2362 */
static void block_correction(orc *s, double *phi) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    double *bl = zalloc(ni), *blp = zalloc(ni), *blm = zalloc(ni), *blc = zalloc(ni);
    double *dphi = zalloc(ni), *pib = zalloc(ni), *qib = zalloc(ni);
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j)
            for (int i = 1; i < s->nim1; ++i) {
                size_t c = IX(s, i, j, k);
                bl[i] += s->ap[c] - s->an[c] - s->as[c] - s->at[c] - s->ab[c];
                blp[i] += s->ae[c];
                blm[i] += s->aw[c];
                blc[i] += s->ae[c] * phi[c + 1] + s->aw[c] * phi[c - 1] + s->an[c] * phi[c + sj] +
                          s->as[c] * phi[c - sj] + s->at[c] * phi[c + sk] + s->ab[c] * phi[c - sk] + s->su[c] -
                          s->ap[c] * phi[c];
            }
    pib[1] = blp[1] / bl[1];
    qib[1] = blc[1] / bl[1];
    for (int i = 2; i < s->nim1; ++i) {
        double den = bl[i] - blm[i] * pib[i - 1];
        pib[i] = blp[i] / den;
        qib[i] = (blc[i] + blm[i] * qib[i - 1]) / den;
    }
    dphi[s->nim2] = qib[s->nim2];
    for (int i = s->nim2 - 1; i > 0; --i) dphi[i] = pib[i] * dphi[i + 1] + qib[i];
    for (int k = 1; k < s->nkm1; ++k)
        for (int j = 1; j < s->njm1; ++j)
            for (int i = 1; i < s->nim1; ++i) phi[IX(s, i, j, k)] += dphi[i];
    free(bl); free(blp); free(blm); free(blc); free(dphi); free(pib); free(qib);
}

/* Reference-form Thomas for one x-line (pr/qr recurrences of
 * CFDSolverManager.cpp:2501-2521); used by the unit tests. */
void orc_thomas(int n, const double *aw, const double *ap, const double *ae, const double *d, double *x) {
    double *pr = zalloc(n + 2), *qr = zalloc(n + 2);
    pr[0] = 0.0;
    qr[0] = x[0];
    for (int i = 1; i <= n; ++i) {
        double den = ap[i] - aw[i] * pr[i - 1];
        pr[i] = ae[i] / den;
        qr[i] = (d[i] + aw[i] * qr[i - 1]) / den;
    }
    for (int i = n; i > 0; --i) x[i] = pr[i] * x[i + 1] + qr[i];
    free(pr);
    free(qr);
}

/* Gauss-Seidel-ordered x-line sweeps: for ks in 0,1; k = kend..kbeg; js in 0,1;
 * j = jbeg..jend; Thomas along i in [ibeg, iend] with phi[ibeg-1] and
 * phi[iend+1] fixed.  Whole-domain form: CFDSolverManager.cpp:2459-2526;
 * melt-pool form with the |denom|<=1e-12 guard: CFDSolverManager.cpp:2367-2454. */
static void line_sweeps(orc *s, double *phi, int ibeg, int iend, int jbeg, int jend, int kbeg, int kend,
                        int guard) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    double *pr = zalloc(ni + 1), *qr = zalloc(ni + 1);
    for (int ks = 0; ks < 2; ++ks)
        for (int k = kend; k >= kbeg; --k)
            for (int js = 0; js < 2; ++js)
                for (int j = jbeg; j <= jend; ++j) {
                    size_t row = IX(s, 0, j, k);
                    pr[ibeg - 1] = 0.0;
                    qr[ibeg - 1] = phi[row + ibeg - 1];
                    for (int i = ibeg; i <= iend; ++i) {
                        size_t c = row + i;
                        double d = s->at[c] * phi[c + sk] + s->ab[c] * phi[c - sk] + s->an[c] * phi[c + sj] +
                                   s->as[c] * phi[c - sj] + s->su[c];
                        double den = s->ap[c] - s->aw[c] * pr[i - 1];
                        if (guard && fabs(den) <= 1e-12) den += SMALL;
                        pr[i] = s->ae[c] / den;
                        qr[i] = (d + s->aw[c] * qr[i - 1]) / den;
                    }
                    for (int i = iend; i >= ibeg; --i) phi[row + i] = pr[i] * phi[row + i + 1] + qr[i];
                }
    free(pr);
    free(qr);
}

/* enthalpy -> temperature and liquid fraction, CFDSolverManager.cpp:2531-2571 */
double orc_h_to_t(const amcfd_params *p, double h, double *fl) {
    if (h >= p->hlcal) {
        *fl = 1.0;
        return (h - p->hlcal) / p->acpl + p->tliquid;
    } else if (h <= p->hsmelt) {
        *fl = 0.0;
        return (sqrt(p->acpb * p->acpb + 2 * p->acpa * h) - p->acpb) / p->acpa;
    }
    *fl = (h - p->hsmelt) / (p->hlcal - p->hsmelt);
    return (p->tliquid - p->tsolid) * (*fl) + p->tsolid;
}

static void enthalpy_to_temp(orc *s) {
    size_t nn = (size_t)s->ni * s->nj * s->nk;
    for (size_t c = 0; c < nn; ++c) s->temp[c] = orc_h_to_t(&s->p, s->enthalpy[c], &s->fracl[c]);
}

/* peak temperature, pool extents and momentum box, BoundCondManager.cpp:280-447 */
static void melt_pool_shape(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const int nim1 = s->nim1, njm1 = s->njm1, nkm1 = s->nkm1, nim2 = s->nim2, njm2 = s->njm2,
              nkm2 = s->nkm2;
    const double ts = s->p.tsolid;
    const double *T = s->temp;
    size_t nn = (size_t)ni * nj * s->nk;
    double tp = T[0];
    for (size_t c = 1; c < nn; ++c)
        if (tp < T[c]) tp = T[c];
    s->tpeak = tp;
    if (s->tpeak <= ts) {
        s->alen = 0.0;
        s->depth = 0.0;
        s->width = 0.0;
        return;
    }
#define TT(i, j, k) T[IX(s, (i), (j), (k))]
    const int js = s->jstart;
    /* length along the top ghost row through the beam */
    s->imax = s->istart;
    s->imin = s->istart;
    s->alen = 0.0;
    for (int i = s->istart; i < nim1; ++i) {
        s->imax = i;
        if (TT(i, js, nkm1) <= ts) break;
    }
    double dtdx = (s->x[s->imax] - s->x[s->imax + 1]) / (TT(s->imax, js, nkm1) - TT(s->imax + 1, js, nkm1));
    double xxmax = s->x[s->imax] + (ts - TT(s->imax, js, nkm1)) * dtdx;
    for (int i = s->istart; i > 0; --i) {
        s->imin = i;
        if (TT(i, js, nkm1) < ts) break;
    }
    dtdx = (s->x[s->imin] - s->x[s->imin - 1]) / (TT(s->imin, js, nkm1) - TT(s->imin - 1, js, nkm1));
    double xxmin = s->x[s->imin] + (ts - TT(s->imin, js, nkm1)) * dtdx;
    s->alen = xxmax - xxmin;
    /* depth in the jstart slice */
    s->kmin = nkm2;
    s->depth = 0.0;
    for (int i = 1; i < nim1; ++i)
        for (int k = nkm2; k > 0; --k) {
            if (TT(i, js, k) < ts) break;
            s->kmin = imn(s->kmin, k);
        }
    s->kmin--;
    if (s->kmin == 0) {
        s->depth = s->z[nkm1] - s->z[0];
    } else {
        for (int i = 1; i < nim1; ++i) {
            if (TT(i, js, s->kmin + 1) < ts) continue;
            double dtdz = (s->z[s->kmin] - s->z[s->kmin - 1]) / (TT(i, js, s->kmin) - TT(i, js, s->kmin - 1));
            double dep = s->z[nkm1] - s->z[s->kmin] + (TT(i, js, s->kmin) - ts) * dtdz;
            s->depth = mx(dep, s->depth);
        }
    }
    s->kmax = nkm2;
    /* width on the top ghost plane */
    s->jmax = js;
    s->jmin = js;
    s->width = 0.0;
    for (int i = 1; i < nim1; ++i)
        for (int j = js; j < njm1; ++j) {
            if (TT(i, j, nkm1) < ts) break;
            s->jmax = imx(s->jmax, j);
        }
    s->jmax++;
    double yymax = s->y[0];
    if (s->jmax != js) {
        for (int i = 1; i < nim1; ++i) {
            if (TT(i, s->jmax - 1, nkm1) < ts) continue;
            double dtdy = (s->y[s->jmax] - s->y[s->jmax + 1]) / (TT(i, s->jmax, nkm1) - TT(i, s->jmax + 1, nkm1));
            double wid = s->y[s->jmax] + (ts - TT(i, s->jmax, nkm1)) * dtdy;
            yymax = mx(wid, yymax);
        }
    }
    for (int i = 1; i < nim1; ++i)
        for (int j = js; j > 0; --j) {
            if (TT(i, j, nkm1) < ts) break;
            s->jmin = imn(s->jmin, j);
        }
    s->jmin--;
    double yymin = s->y[js];
    if (s->jmin != js) {
        for (int i = 1; i < nim1; ++i) {
            if (TT(i, s->jmin + 1, nkm1) < ts) continue;
            double dtdy = (s->y[s->jmin] - s->y[s->jmin - 1]) / (TT(i, s->jmin, nkm1) - TT(i, s->jmin - 1, nkm1));
            double wid = s->y[s->jmin] + (ts - TT(i, s->jmin, nkm1)) * dtdy;
            yymin = mn(wid, yymin);
        }
    }
    s->width = yymax - yymin;
#undef TT
    s->istat = imx(s->imin - 3, 1);
    s->iend = imn(s->imax + 3, nim2);
    s->jstat = imx(s->jmin - 3, 1);
    s->jend = imn(s->jmax + 2, njm2);
    s->kstat = imx(s->kmin - 2, 2);
    s->istatp1 = s->istat + 1;
    s->iendm1 = s->iend - 1;
}

/* zero face velocities next to solid cells, CFDSolverManager.cpp:2577-2620 */
static void clean_velocities(orc *s) {
    const double ts = s->p.tsolid;
    const size_t sj = (size_t)s->ni, sk = (size_t)s->ni * s->nj;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                if (mn(s->temp[c], s->temp[c + 1]) <= ts) s->uvel[c + 1] = 0.0;
                if (mn(s->temp[c], s->temp[c + sj]) <= ts) s->vvel[c + sj] = 0.0;
                if (mn(s->temp[c], s->temp[c + sk]) <= ts) s->wvel[c + sk] = 0.0;
            }
}

/* ------------------------------------------------------------------------ */
/* momentum coefficients on the staggered grid,                              */
/* CFDSolverManager.cpp:771-1111                                             */
/* ------------------------------------------------------------------------ */
/* p*q / (wq*q + wp*p): the distance-weighted harmonic mean used for every
 * face viscosity; the caller passes the weights in the reference's order. */
static inline double hmean(double p, double q, double wq, double wp) { return p * q / (wq * q + wp * p); }

static void discretize_u(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    const double delt = s->p.delt;
    const double *V = s->vis, *U = s->uvel, *Vv = s->vvel, *W = s->wvel;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                double fx = s->fracx[i - 1];
                double vn = Vv[c + sj] * (1.0 - fx) + Vv[c + sj - 1] * fx;
                double vs = Vv[c] * (1.0 - fx) + Vv[c - 1] * fx;
                double ue = (U[c + 1] + U[c]) * 0.5;
                double uw = (U[c - 1] + U[c]) * 0.5;
                double wt = W[c + sk] * (1.0 - fx) + W[c + sk - 1] * fx;
                double wb = W[c] * (1.0 - fx) + W[c - 1] * fx;
                double r = s->rho[c];
                double aik = s->areauik[k * ni + i], ajk = s->areajk[k * nj + j], aij = s->areauij[j * ni + i];
                double fn = vn * r * aik, fs = vs * r * aik, fe = ue * r * ajk, fw = uw * r * ajk;
                double ft = wt * r * aij, fb = wb * r * aij;
                double visu = hmean(V[c], V[c - 1], fx, 1.0 - fx);
                double visun = hmean(V[c + sj], V[c + sj - 1], fx, 1.0 - fx);
                double visus = hmean(V[c - sj], V[c - sj - 1], fx, 1.0 - fx);
                double visn = (j == s->njm2) ? visu : hmean(visu, visun, 1.0 - s->fracy[j], s->fracy[j]);
                double viss = (j == 1) ? visu : hmean(visu, visus, s->fracy[j - 1], 1.0 - s->fracy[j - 1]);
                double vise = V[c], visw = V[c - 1];
                double visut = hmean(V[c + sk], V[c + sk - 1], fx, 1.0 - fx);
                double visub = hmean(V[c - sk], V[c - sk - 1], fx, 1.0 - fx);
                double vist = (k == s->nkm2) ? visu : hmean(visut, visu, s->fracz[k], 1.0 - s->fracz[k]);
                double visb = (k == 1) ? visu : hmean(visub, visu, 1.0 - s->fracz[k - 1], s->fracz[k - 1]);
                double dn = visn * aik * s->dypsinv[j + 1];
                double ds = viss * aik * s->dypsinv[j];
                double de = vise * ajk / (s->xu[i + 1] - s->xu[i]);
                double dw = visw * ajk / (s->xu[i] - s->xu[i - 1]);
                double dt = vist * aij * s->dzpbinv[k + 1];
                double db = visb * aij * s->dzpbinv[k];
                s->an[c] = coef_neg(s, dn, fn);
                s->as[c] = coef_pos(s, ds, fs);
                s->ae[c] = coef_neg(s, de, fe);
                s->aw[c] = coef_pos(s, dw, fw);
                s->at[c] = coef_neg(s, dt, ft);
                s->ab[c] = coef_pos(s, db, fb);
                s->apnot[c] = r * s->volume_u[c] / delt;
                double delf = fn - fs + fe - fw + ft - fb;
                double cp0 = mx(0.0, delf), cp1 = mn(0.0, delf);
                s->su[c] = -cp1 * U[c];
                s->su[c] += ajk * (s->pressure[c - 1] - s->pressure[c]);
                s->sp[c] = -cp0;
                s->su[c] += s->apnot[c] * s->unot[c];
                double dudxp = (U[c + 1] - U[c]) / (s->xu[i + 1] - s->xu[i]);
                double dudxm = (U[c] - U[c - 1]) / (s->xu[i] - s->xu[i - 1]);
                s->su[c] += (vise * dudxp - visw * dudxm) * ajk;
                double dvdxp = (Vv[c + sj] - Vv[c + sj - 1]) * s->dxpwinv[i];
                double dvdxm = (Vv[c] - Vv[c - 1]) * s->dxpwinv[i];
                s->su[c] += (visn * dvdxp - viss * dvdxm) * aik;
                double dwdxp = (W[c + sk] - W[c + sk - 1]) * s->dxpwinv[i];
                double dwdxm = (W[c] - W[c - 1]) * s->dxpwinv[i];
                s->su[c] += (vist * dwdxp - visb * dwdxm) * aij;
            }
}

static void discretize_v(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    const double delt = s->p.delt;
    const double *V = s->vis, *U = s->uvel, *Vv = s->vvel, *W = s->wvel;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                double fy = s->fracy[j - 1];
                double vn = (Vv[c] + Vv[c + sj]) * 0.5;
                double vs = (Vv[c] + Vv[c - sj]) * 0.5;
                double ue = U[c + 1] * (1.0 - fy) + U[c - sj + 1] * fy;
                double uw = U[c] * (1.0 - fy) + U[c - sj] * fy;
                double wt = W[c + sk] * (1.0 - fy) + W[c + sk - sj] * fy;
                double wb = W[c] * (1.0 - fy) + W[c - sj] * fy;
                double r = s->rho[c];
                double aik = s->areaik[k * ni + i], ajk = s->areavjk[k * nj + j], aij = s->areavij[j * ni + i];
                double fn = vn * r * aik, fs = vs * r * aik, fe = ue * r * ajk, fw = uw * r * ajk;
                double ft = wt * r * aij, fb = wb * r * aij;
                double visn = V[c], viss = V[c - sj];
                double visv = hmean(V[c], V[c - sj], fy, 1.0 - fy);
                double visve = hmean(V[c + 1], V[c - sj + 1], fy, 1.0 - fy);
                double visvw = hmean(V[c - 1], V[c - sj - 1], fy, 1.0 - fy);
                double vise = (i == s->nim2) ? visv : hmean(visv, visve, 1.0 - s->fracx[i], s->fracx[i]);
                double visw = (i == 1) ? visv : hmean(visv, visvw, s->fracx[i - 1], 1.0 - s->fracx[i - 1]);
                double visvt = hmean(V[c + sk], V[c + sk - sj], fy, 1.0 - fy);
                double visvb = hmean(V[c - sk], V[c - sk - sj], fy, 1.0 - fy);
                double vist = (k == s->nkm2) ? visv : hmean(visv, visvt, 1.0 - s->fracz[k], s->fracz[k]);
                double visb = (k == 1) ? visv : hmean(visv, visvb, s->fracz[k - 1], 1.0 - s->fracz[k - 1]);
                double dn = visn * aik / (s->yv[j + 1] - s->yv[j]);
                double ds = viss * aik / (s->yv[j] - s->yv[j - 1]);
                double de = vise * ajk * s->dxpwinv[i + 1];
                double dw = visw * ajk * s->dxpwinv[i];
                double dt = vist * aij * s->dzpbinv[k + 1];
                double db = visb * aij * s->dzpbinv[k];
                s->an[c] = coef_neg(s, dn, fn);
                s->as[c] = coef_pos(s, ds, fs);
                s->ae[c] = coef_neg(s, de, fe);
                s->aw[c] = coef_pos(s, dw, fw);
                s->at[c] = coef_neg(s, dt, ft);
                s->ab[c] = coef_pos(s, db, fb);
                s->apnot[c] = r * s->volume_v[c] / delt;
                double delf = fn - fs + fe - fw + ft - fb;
                double cp0 = mx(0.0, delf), cp1 = mn(0.0, delf);
                s->su[c] = -cp1 * Vv[c];
                s->su[c] += aik * (s->pressure[c - sj] - s->pressure[c]);
                s->sp[c] = -cp0;
                s->su[c] += s->apnot[c] * s->vnot[c];
                double dudyp = (U[c + 1] - U[c - sj + 1]) * s->dypsinv[j];
                double dudym = (U[c] - U[c - sj]) * s->dypsinv[j];
                s->su[c] += (vise * dudyp - visw * dudym) * ajk;
                double dvdyp = (Vv[c + sj] - Vv[c]) / (s->yv[j + 1] - s->yv[j]);
                double dvdym = (Vv[c] - Vv[c - sj]) / (s->yv[j] - s->yv[j - 1]);
                s->su[c] += (visn * dvdyp - viss * dvdym) * aik;
                double dwdyp = (W[c + sk] - W[c + sk - sj]) * s->dypsinv[j];
                double dwdym = (W[c] - W[c - sj]) * s->dypsinv[j];
                s->su[c] += (vist * dwdyp - visb * dwdym) * aij;
            }
}

static void discretize_w(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    const double delt = s->p.delt;
    const double *V = s->vis, *U = s->uvel, *Vv = s->vvel, *W = s->wvel;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                double fz = s->fracz[k - 1];
                double vn = Vv[c + sj] * (1.0 - fz) + Vv[c - sk + sj] * fz;
                double vs = Vv[c] * (1.0 - fz) + Vv[c - sk] * fz;
                double ue = U[c + 1] * (1.0 - fz) + U[c - sk + 1] * fz;
                double uw = U[c] * (1.0 - fz) + U[c - sk] * fz;
                double wt = (W[c] + W[c + sk]) * 0.5;
                double wb = (W[c] + W[c - sk]) * 0.5;
                double r = s->rho[c];
                double aik = s->areawik[k * ni + i], ajk = s->areawjk[k * nj + j], aij = s->areaij[j * ni + i];
                double fn = vn * r * aik, fs = vs * r * aik, fe = ue * r * ajk, fw = uw * r * ajk;
                double ft = wt * r * aij, fb = wb * r * aij;
                double visw0 = hmean(V[c], V[c - sk], fz, 1.0 - fz);
                double viswn = hmean(V[c + sj], V[c - sk + sj], fz, 1.0 - fz);
                double visws = hmean(V[c - sj], V[c - sk - sj], fz, 1.0 - fz);
                double visn = (j == s->njm2) ? visw0 : hmean(visw0, viswn, 1.0 - s->fracy[j], s->fracy[j]);
                double viss = (j == 1) ? visw0 : hmean(visw0, visws, s->fracy[j - 1], 1.0 - s->fracy[j - 1]);
                double viswe = hmean(V[c + 1], V[c - sk + 1], fz, 1.0 - fz);
                double visww = hmean(V[c - 1], V[c - sk - 1], fz, 1.0 - fz);
                double vise = (i == s->nim2) ? visw0 : hmean(visw0, viswe, 1.0 - s->fracx[i], s->fracx[i]);
                double visw = (i == 1) ? visw0 : hmean(visw0, visww, s->fracx[i - 1], 1.0 - s->fracx[i - 1]);
                double vist = V[c], visb = V[c - sk];
                double dn = visn * aik * s->dypsinv[j + 1];
                double ds = viss * aik * s->dypsinv[j];
                double de = vise * ajk * s->dxpwinv[i + 1];
                double dw = visw * ajk * s->dxpwinv[i];
                double dt = vist * aij / (s->zw[k + 1] - s->zw[k]);
                double db = visb * aij / (s->zw[k] - s->zw[k - 1]);
                s->an[c] = coef_neg(s, dn, fn);
                s->as[c] = coef_pos(s, ds, fs);
                s->ae[c] = coef_neg(s, de, fe);
                s->aw[c] = coef_pos(s, dw, fw);
                s->at[c] = coef_neg(s, dt, ft);
                s->ab[c] = coef_pos(s, db, fb);
                s->apnot[c] = r * s->volume_w[c] / delt;
                double delf = fn - fs + fe - fw + ft - fb;
                double cp0 = mx(0.0, delf), cp1 = mn(0.0, delf);
                s->su[c] = -cp1 * W[c];
                s->su[c] += aij * (s->pressure[c - sk] - s->pressure[c]);
                s->sp[c] = -cp0;
                s->su[c] += s->apnot[c] * s->wnot[c];
                double dudzp = (U[c + 1] - U[c - sk + 1]) * s->dzpbinv[k];
                double dudzm = (U[c] - U[c - sk]) * s->dzpbinv[k];
                s->su[c] += (vise * dudzp - visw * dudzm) * ajk;
                double dvdzp = (Vv[c + sj] - Vv[c - sk + sj]) * s->dzpbinv[k];
                double dvdzm = (Vv[c] - Vv[c - sk]) * s->dzpbinv[k];
                s->su[c] += (visn * dvdzp - viss * dvdzm) * aik;
                double dwdzp = (W[c + sk] - W[c]) / (s->zw[k + 1] - s->zw[k]);
                double dwdzm = (W[c] - W[c - sk]) / (s->zw[k] - s->zw[k - 1]);
                s->su[c] += (vist * dwdzp - visb * dwdzm) * aij;
            }
}

/* pressure-correction coefficients + mass imbalance, CFDSolverManager.cpp:1112-1156 */
static void discretize_p(orc *s) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    s->resorm = 0.0;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                double r = s->rho[c];
                double aik = s->areaik[k * ni + i], ajk = s->areajk[k * nj + j], aij = s->areaij[j * ni + i];
                s->an[c] = aik * s->dvy[c + sj] * r;
                s->as[c] = aik * s->dvy[c] * r;
                s->ae[c] = ajk * s->dux[c + 1] * r;
                s->aw[c] = ajk * s->dux[c] * r;
                s->at[c] = aij * s->dwz[c + sk] * r;
                s->ab[c] = aij * s->dwz[c] * r;
                double fn = s->vvel[c + sj] * aik * r, fs = s->vvel[c] * aik * r;
                double fe = s->uvel[c + 1] * ajk * r, fw = s->uvel[c] * ajk * r;
                double ft = s->wvel[c + sk] * aij * r, fb = s->wvel[c] * aij * r;
                double delf = fn - fs + fe - fw + ft - fb;
                s->sp[c] = 0.0;
                s->su[c] = -delf;
                s->resorm += fabs(delf);
            }
}

/* Darcy / buoyancy sources, top fold, ap, d-coefficients, relaxation and
 * solid pinning for one velocity component, CFDSolverManager.cpp:1408-1644 */
static void source_velocity(orc *s, int ivar) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    const double ts = p->tsolid;
    double *phi = ivar == 1 ? s->uvel : ivar == 2 ? s->vvel : s->wvel;
    double *dcoef = ivar == 1 ? s->dux : ivar == 2 ? s->dvy : s->dwz;
    const double *vol = ivar == 1 ? s->volume_u : ivar == 2 ? s->volume_v : s->volume_w;
    const double urf = ivar == 1 ? p->urfu : ivar == 2 ? p->urfv : p->urfw;
    const size_t back = ivar == 1 ? 1 : ivar == 2 ? sj : sk;   /* upstream node of the face */
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                double w = ivar == 1 ? s->fracx[i - 1] : ivar == 2 ? s->fracy[j - 1] : s->fracz[k - 1];
                double flf = s->fracl[c] * (1.0 - w) + s->fracl[c - back] * w;
                double rhoscan = s->rho[c] * s->scanvel;
                double boufac = s->rho[c] * p->grav * p->beta;
                if (flf > 0) {
                    double term = p->darcyKo * ((1.0 - flf) * (1.0 - flf)) / ((flf * flf * flf) + SMALL);
                    s->sp[c] -= term * vol[c];
                    if (p->steady) {
                        double a = ivar == 1 ? s->areajk[k * nj + j] : ivar == 2 ? s->areavjk[k * nj + j]
                                                                                 : s->areawjk[k * nj + j];
                        s->su[c] += rhoscan * a * (phi[c - 1] - phi[c]);
                    }
                    if (ivar == 3) {
                        double tw = s->temp[c] * (1.0 - w) + s->temp[c - sk] * w;
                        s->su[c] += boufac * vol[c] * (tw - ts);
                    }
                }
            }
    if (ivar != 3) {
        /* top ghost folded into the source, CFDSolverManager.cpp:1440-1449 */
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, s->nkm2);
                s->su[c] += s->at[c] * phi[c + sk];
                s->sp[c] -= s->at[c];
                s->at[c] = 0.0;
            }
    }
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                s->ap[c] = s->an[c] + s->as[c] + s->ae[c] + s->aw[c] + s->at[c] + s->ab[c] + s->apnot[c] - s->sp[c];
                double area = ivar == 1 ? s->areajk[k * nj + j] : ivar == 2 ? s->areaik[k * ni + i]
                                                                            : s->areaij[j * ni + i];
                dcoef[c] = area / s->ap[c];
                s->ap[c] /= urf;
                s->su[c] += (1.0 - urf) * s->ap[c] * phi[c];
                dcoef[c] *= urf;
                if (mn(s->temp[c], s->temp[c - back]) <= ts) {
                    s->su[c] = 0.0;
                    s->an[c] = 0.0;
                    s->as[c] = 0.0;
                    s->ae[c] = 0.0;
                    s->aw[c] = 0.0;
                    s->at[c] = 0.0;
                    s->ab[c] = 0.0;
                    s->ap[c] = LARGE;
                }
            }
}

static void source_pressure(orc *s) {
    const double ts = s->p.tsolid;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                s->ap[c] = s->an[c] + s->as[c] + s->ae[c] + s->aw[c] + s->at[c] + s->ab[c] - s->sp[c];
                if (s->temp[c] <= ts) {
                    s->su[c] = 0.0;
                    s->an[c] = 0.0;
                    s->as[c] = 0.0;
                    s->ae[c] = 0.0;
                    s->aw[c] = 0.0;
                    s->at[c] = 0.0;
                    s->ab[c] = 0.0;
                    s->ap[c] = LARGE;
                }
            }
}

/* momentum and continuity residuals, CFDSolverManager.cpp:1960-2072 */
static void residual_flow(orc *s, int ivar) {
    const int ni = s->ni, nj = s->nj;
    const size_t sj = (size_t)ni, sk = (size_t)ni * nj;
    if (ivar == 4) {
        double denom = 0.0;
        for (int k = s->kstat; k < s->nkm1; ++k)
            for (int j = s->jstat; j <= s->jend; ++j)
                for (int i = s->istatp1; i <= s->iendm1; ++i) {
                    size_t c = IX(s, i, j, k);
                    double dtp = (fabs(s->uvel[c]) + fabs(s->uvel[c + 1])) * s->areajk[k * nj + j] +
                                 (fabs(s->vvel[c]) + fabs(s->vvel[c + sj])) * s->areaik[k * ni + i] +
                                 (fabs(s->wvel[c]) + fabs(s->wvel[c + sk])) * s->areaij[j * ni + i];
                    denom += 0.5 * fabs(dtp);
                }
        denom *= s->p.dens;
        s->resorm /= (denom + SMALL);
        return;
    }
    const double *phi = ivar == 1 ? s->uvel : ivar == 2 ? s->vvel : s->wvel;
    double sumd = 0.0;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                double r = s->an[c] * phi[c + sj] + s->as[c] * phi[c - sj] + s->ae[c] * phi[c + 1] +
                           s->aw[c] * phi[c - 1] + s->at[c] * phi[c + sk] + s->ab[c] * phi[c - sk] + s->su[c] -
                           s->ap[c] * phi[c];
                sumd += fabs(r);
            }
    if (ivar == 1) {
        double umaxt = 0.0;
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) umaxt = mx(umaxt, fabs(s->uvel[IX(s, i, j, s->nkm1)]));
        s->refmom = 0.25 * M_PI * (s->width * s->width) * s->p.dens * (umaxt * umaxt);
        s->resoru = sumd / s->refmom;
    } else if (ivar == 2) {
        s->resorv = sumd / s->refmom;
    } else {
        s->resorw = sumd / s->refmom;
    }
}

/* velocity and pressure correction, CFDSolverManager.cpp:2692-2758 */
static void correct_pressure(orc *s) {
    const double ts = s->p.tsolid;
    const size_t sj = (size_t)s->ni, sk = (size_t)s->ni * s->nj;
    const double *T = s->temp, *pp = s->pp;
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                if (mn(T[c], T[c - 1]) > ts) s->uvel[c] += s->dux[c] * (pp[c - 1] - pp[c]);
                if (mn(T[c], T[c - sj]) > ts) s->vvel[c] += s->dvy[c] * (pp[c - sj] - pp[c]);
                if (mn(T[c], T[c - sk]) > ts) s->wvel[c] += s->dwz[c] * (pp[c - sk] - pp[c]);
            }
    for (int k = s->kstat; k < s->nkm1; ++k)
        for (int j = s->jstat; j <= s->jend; ++j)
            for (int i = s->istatp1; i <= s->iendm1; ++i) {
                size_t c = IX(s, i, j, k);
                if (T[c] > ts) {
                    s->pressure[c] += s->p.urfp * s->pp[c];
                    s->pp[c] = 0.0;
                }
            }
}

/* boundary heat fluxes and energy balance, BoundCondManager.cpp:746-858 */
static void heat_fluxes(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj;
    const int nim1 = s->nim1, njm1 = s->njm1, nkm1 = s->nkm1, nim2 = s->nim2, njm2 = s->njm2,
              nkm2 = s->nkm2;
    const double *h = s->enthalpy, *D = s->diff;
    s->fluxwest = s->fluxeast = s->fluxbottom = s->fluxtop = s->fluxnorth = s->fluxsouth = 0.0;
    s->accul = 0.0;
    s->heatvol = 0.0;
    for (int k = 1; k < nkm1; ++k)
        for (int j = 1; j < njm1; ++j) {
            double fi = D[IX(s, 1, j, k)] * (h[IX(s, 0, j, k)] - h[IX(s, 1, j, k)]) * s->dxpwinv[1];
            double fl = D[IX(s, nim2, j, k)] * (h[IX(s, nim1, j, k)] - h[IX(s, nim2, j, k)]) * s->dxpwinv[nim1];
            s->fluxwest += s->areajk[k * nj + j] * fi;
            s->fluxeast += s->areajk[k * nj + j] * fl;
        }
    for (int j = 1; j < njm1; ++j)
        for (int i = 1; i < nim1; ++i) {
            double fk = D[IX(s, i, j, 1)] * (h[IX(s, i, j, 0)] - h[IX(s, i, j, 1)]) * s->dzpbinv[1];
            double fn = D[IX(s, i, j, nkm2)] * (h[IX(s, i, j, nkm1)] - h[IX(s, i, j, nkm2)]) * s->dzpbinv[nkm1];
            s->fluxbottom += s->areaij[j * ni + i] * fk;
            s->fluxtop += s->areaij[j * ni + i] * fn;
        }
    for (int k = 1; k < nkm1; ++k)
        for (int i = 1; i < nim1; ++i) {
            double fj = D[IX(s, i, 1, k)] * (h[IX(s, i, 0, k)] - h[IX(s, i, 1, k)]) * s->dypsinv[1];
            double fm = D[IX(s, i, njm2, k)] * (h[IX(s, i, njm1, k)] - h[IX(s, i, njm2, k)]) * s->dypsinv[njm1];
            s->fluxsouth += s->areaik[k * ni + i] * fj;
            s->fluxnorth += s->areaik[k * ni + i] * fm;
        }
    if (!p->steady) {
        for (int k = 1; k < nkm1; ++k)
            for (int j = 1; j < njm1; ++j)
                for (int i = 1; i < nim1; ++i) {
                    size_t c = IX(s, i, j, k);
                    double dh = h[c] - s->hnot[c] + (s->fracl[c] - s->fraclnot[c]) * p->hlatnt;
                    s->accul += s->volume[c] * s->rho[c] * dh / p->delt;
                    s->heatvol += s->sourceinput[c] * s->volume[c];
                }
    }
    s->heatout = (s->fluxnorth + s->fluxbottom + s->fluxwest + s->fluxeast + s->fluxsouth) - s->ahtoploss;
    if (p->steady) s->ratio = -s->heatout / s->heatinlaser;
    else s->ratio = (s->heatvol + s->heatinlaser) / (s->accul - s->heatout);
}

/* ------------------------------------------------------------------------ */
/* SIMPLE loop, CFDSolverManager.cpp:595-722                                 */
/* ------------------------------------------------------------------------ */
void orc_energy_iteration(orc *s) {
    apply_bcs(s, 5);
    discretize_energy(s);
    source_energy(s);
    residual_energy(s);
    material_props(s);
    block_correction(s, s->enthalpy);
    line_sweeps(s, s->enthalpy, 1, s->nim2, 1, s->njm2, 1, s->nkm2, 0);
    enthalpy_to_temp(s);
    melt_pool_shape(s);
}

static void flow_iteration(orc *s) {
    clean_velocities(s);
    for (int ivar = 1; ivar < 5; ++ivar) {
        apply_bcs(s, ivar);
        switch (ivar) {
        case 1: discretize_u(s); break;
        case 2: discretize_v(s); break;
        case 3: discretize_w(s); break;
        case 4: discretize_p(s); break;
        }
        if (ivar < 4) source_velocity(s, ivar);
        else source_pressure(s);
        residual_flow(s, ivar);
        double *phi = ivar == 1 ? s->uvel : ivar == 2 ? s->vvel : ivar == 3 ? s->wvel : s->pp;
        line_sweeps(s, phi, s->istatp1, s->iendm1, s->jstat, s->jend, s->kstat, s->nkm2, 1);
        if (ivar == 4) correct_pressure(s);
    }
}

/* species conservation after the flow, CFDSolverManager.cpp:677-686 */
static void species_iteration(orc *s) {
    bc_species(s);
    discretize_species(s);
    source_species(s);
    residual_species(s);
    block_correction(s, s->conc);
    line_sweeps(s, s->conc, 1, s->nim2, 1, s->njm2, 1, s->nkm2, 0);
}

/* Temperature gradient G and growth rate R, CFDSolverManager.cpp:2763-2804:
 * in every x-row (k < nkm1, all j) at the first node i with T >= Ts, written
 * to node i-1 from the flat-index neighbours i-2 / i, +-one row, +-one plane
 * of i-1 (so rows at the edges read into the adjacent row or plane, as the
 * reference's flat arrays do).  Where the reference reads its coordinate or
 * temperature arrays outside their bounds (first such node at i = 0, or in
 * row j = 0 or plane k = 0) G and R are set to NaN. */
static void solidification(orc *s) {
    const amcfd_params *p = &s->p;
    if (s->tpeak <= p->tsolid) return;
    const long ni = s->ni, nj = s->nj;
    const long sj = ni, sk = ni * nj;
    for (int k = 0; k < s->nkm1; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                long c = ((long)k * nj + j) * ni + i;
                if (!(s->temp[c] >= p->tsolid)) continue;
                long o = c - 1;
                if (i == 0 || j == 0 || k == 0) {
                    if (o >= 0) {
                        s->grad[o] = NAN;
                        s->rate[o] = NAN;
                    }
                    break;
                }
                double gx = (s->temp[c] - s->temp[c - 2]) / (s->x[i] - s->x[i - 1]) / 2;
                double gy = (s->temp[o + sj] - s->temp[o - sj]) / (s->y[j] - s->y[j - 1]) / 2;
                double gz = (s->temp[o + sk] - s->temp[o - sk]) / (s->z[k] - s->z[k - 1]) / 2;
                s->grad[o] = sqrt(gx * gx + gy * gy + gz * gz);
                s->rate[o] = s->scanvel * gx / s->grad[o];
                break;
            }
}

/* ------------------------------------------------------------------------ */
/* energy-based free surface, FreeSurfaceManager.cpp:72-342 (output times)   */
/* ------------------------------------------------------------------------ */
/* surface shift for one Lagrange multiplier: relaxed fixed-point iteration
 * of the surface energy balance over the molten top cells, :251-341.  The
 * largest relative change is kept across iterations (it is not reset), so
 * once it is infinite (first iteration from zero) the loop runs to its
 * limit. */
static void fs_ebal(orc *s, double alambda) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj, nkm1 = s->nkm1;
    const int nitermax = (int)p->maxiterenergy;
    const double c1 = p->dens * p->grav, gamma = p->surftens, urfree = p->urfree, tol = p->tolenergy;
    const double *xu = s->xu, *yv = s->yv, *x = s->x, *y = s->y;
    double *fi = s->fi, *zzm = s->zzm;
    int niter = 0;
    double difmax = 0.0;
    for (int c = 0; c < ni * nj; ++c) zzm[c] = 0.0;
    while (niter < nitermax) {
        niter++;
        double pmax = -INFINITY;   /* the OpenMP max-reduction's private copy */
        for (int j = 1; j < s->njm1; ++j)
            for (int i = 1; i < s->nim1; ++i) {
                if (s->temp[IX(s, i, j, nkm1)] < p->tsolid) continue;
                int c = j * ni + i;
                double dely = yv[j + 1] - yv[j - 1], dely1 = y[j + 1] - y[j - 1];
                double delyp = y[j + 1] - y[j], delym = y[j] - y[j - 1];
                double delx = xu[i + 1] - xu[i - 1], delx1 = x[i + 1] - x[i - 1];
                double delxp = x[i + 1] - x[i], delxm = x[i] - x[i - 1];
                double c2 = s->apdis[c] + alambda;
                double dzdx = (zzm[c + 1] - zzm[c - 1]) / delx1;
                double dzdy = (zzm[c + ni] - zzm[c - ni]) / dely1;
                double dzdyp1 = (zzm[c + ni + 1] - zzm[c - ni + 1]) / dely1;
                double dzdym1 = (zzm[c + ni - 1] - zzm[c - ni - 1]) / dely1;
                double d2zdxdy = (dzdyp1 - dzdym1) / delx1;
                double c3 = pow(1.0 + dzdx * dzdx + dzdy * dzdy, 1.5);
                double c4 = 1.0 + dzdy * dzdy, c5 = 1.0 + dzdx * dzdx;
                double c6 = 2.0 * dzdx * dzdy * d2zdxdy;
                double c7 = zzm[c + 1] / (delxp * delx) + zzm[c - 1] / (delxm * delx);
                double c8 = zzm[c + ni] / (delyp * dely) + zzm[c - ni] / (delym * dely);
                double c9 = 1.0 / (delxp * delx) + 1.0 / (delxm * delx);
                double c10 = 1.0 / (delyp * dely) + 1.0 / (delym * dely);
                double c11 = gamma / c3;
                double term1 = c2 + c11 * (c4 * c7 + c5 * c8 - c6);
                double term2 = c1 + c11 * (c4 * c9 + c5 * c10);
                fi[c] = urfree * term1 / term2 + (1.0 - urfree) * zzm[c];
                double diff1 = fabs((fi[c] - zzm[c]) / zzm[c]);
                pmax = mx(diff1, pmax);
            }
        difmax = (pmax > difmax) ? pmax : difmax;
        if (difmax < tol) break;
        for (int c = 0; c < ni * nj; ++c) zzm[c] = fi[c];
    }
}

/* volume change for one multiplier, :225-246 */
static double fs_delta_vol(orc *s, double alambda) {
    fs_ebal(s, alambda);
    double delv = 0.0;
    for (int j = 1; j < s->njm1; ++j)
        for (int i = 1; i < s->nim1; ++i) delv += s->areaij[j * s->ni + i] * s->fi[j * s->ni + i];
    return delv;
}

/* FreeSurfaceManager::execute, :72-220: bisection on the multiplier so that
 * the surface displaces the added droplet volume, then heights and
 * temperatures of every node column re-interpolated along z.  Returns 1 if
 * the reference would print its bracket warning. */
int orc_free_surface(orc *s) {
    const amcfd_params *p = &s->p;
    const int ni = s->ni, nj = s->nj, nk = s->nk, nkm1 = s->nkm1;
    const int nitermax = (int)p->maxiterbisect;
    double alam1 = p->initlam1, alam2 = p->initlam2, tol = p->tolbisect;
    const double delt = p->delt, startdeltn = 0.0;
    double melttime = delt * (1.0 + startdeltn);
    for (int c = 0; c < ni * nj; ++c) s->apdis[c] = 0.0;
    const double vdp = 2e-9;
    double vd;
    if (s->tpeak <= p->tsolid) {
        vd = 0.0;
        melttime = 0.0;
    } else {
        vd = vdp * (s->timet - melttime - startdeltn * delt);
        if (s->powerindicator < 1) vd = vdp * (p->lasertime - melttime);
    }
    double f = fs_delta_vol(s, alam1) + vd;
    double fmid = fs_delta_vol(s, alam2) + vd;
    int warn = (f * fmid >= 0.0 && f != vd);
    double alamopt, dx, xmid = 0.0;
    if (f < 0.0) {
        alamopt = alam1;
        dx = alam2 - alam1;
    } else {
        alamopt = alam2;
        dx = alam1 - alam2;
    }
    int niter;
    for (niter = 0; niter < nitermax; niter++) {
        dx *= 0.5;
        xmid = alamopt + dx;
        fmid = fs_delta_vol(s, xmid) + vd;
        if (fmid < 0.0) alamopt = xmid;
        if (fabs(dx) < 1.0e-5 || (fabs(fmid) < tol && niter != 0)) break;
    }
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nj; ++j)
            for (int i = 0; i < ni; ++i) {
                size_t c = IX(s, i, j, k);
                double factor = s->fi[j * ni + i] / s->z[nkm1];
                s->zr[c] = s->z[k] * (1.0 - factor);
                if (s->zr[c] > s->z[nkm1]) continue;
                if (s->zr[c] == s->z[nkm1]) s->tempr[c] = s->temp[IX(s, i, j, nkm1)];
                for (int l = nkm1 - 1; l >= 0; l--)
                    if (s->z[l] <= s->zr[c]) {
                        s->tempr[c] = (s->zr[c] - s->z[l]) / (s->z[l + 1] - s->z[l]) *
                                          (s->temp[IX(s, i, j, l + 1)] - s->temp[IX(s, i, j, l)]) +
                                      s->temp[IX(s, i, j, l)];
                        break;
                    }
            }
    s->fs_xmid = xmid;
    s->fs_fmid = fmid;
    s->fs_vd = vd;
    s->fs_niter = niter;
    return warn;
}

void orc_free_surface_monitor(const orc *s, double *xmid, double *fmid, double *fimin, double *fimax,
                              double *vd, int *niter) {
    const int n = s->ni * s->nj;
    double lo = s->fi[0], hi = s->fi[0];
    for (int c = 1; c < n; ++c) {   /* std::min_element / max_element: first extremum */
        if (s->fi[c] < lo) lo = s->fi[c];
        if (hi < s->fi[c]) hi = s->fi[c];
    }
    *xmid = s->fs_xmid;
    *fmid = s->fs_fmid;
    *fimin = lo;
    *fimax = hi;
    *vd = s->fs_vd;
    *niter = s->fs_niter;
}

int orc_solve(orc *s) {
    const amcfd_params *p = &s->p;
    double alenold = s->alen;
    s->niter = 0;
    while (s->niter < p->maxit) {
        s->niter++;
        s->itertot++;
        orc_energy_iteration(s);
        if (s->tpeak > p->tsolid && p->navierstokes) flow_iteration(s);
        if (p->species) species_iteration(s);
        /* max over the flow residuals with std::max_element semantics */
        double res[5] = {s->resorm, s->resoru, s->resorv, s->resorw, s->resorc};
        double amaxres = res[0];
        for (int r = 1; r < 5; ++r)
            if (amaxres < res[r]) amaxres = res[r];
        heat_fluxes(s);
        double ratio = s->ratio;
        if (p->steady) {
            if (ratio <= 1.001 && ratio >= 0.999 && amaxres < p->nonlintol) return 0;
        } else if (s->powerindicator == 1) {
            if (ratio <= 1.01 && ratio >= 0.99 && amaxres < p->nonlintol) return 0;
        } else {
            if (s->resorh < 1e-6 && amaxres < p->nonlintol) return 0;
        }
    }
    if (fabs(s->alen - alenold) < 1.0e-6) s->steadylength = 1;
    return 1;
}

/* end-of-step push-back, CFDSolverManager.cpp:2625-2687 */
void orc_end_step(orc *s) {
    const amcfd_params *p = &s->p;
    if (p->solidification) solidification(s);   /* main.cpp:151-152 */
    if (p->steady) return;
    double track = (double)s->trackindx;
    double trackid = 0.0;
    for (int r = p->n_toolpath; r > 0; --r)
        if (s->timet <= s->tool[6 * (r - 1)]) trackid = s->tool[6 * (r - 1) + 5];
    size_t nn = (size_t)s->ni * s->nj * s->nk;
    for (size_t c = 0; c < nn; ++c) {
        if (s->temp[c] <= p->tsolid) {
            s->uvel[c] = 0.0;
            s->vvel[c] = 0.0;
            s->wvel[c] = 0.0;
        }
        s->unot[c] = s->uvel[c];
        s->vnot[c] = s->vvel[c];
        s->wnot[c] = s->wvel[c];
        s->hnot[c] = s->enthalpy[c];
        s->fraclnot[c] = s->fracl[c];
        if (p->species) {
            s->concnot[c] = s->conc[c];
            s->concl[c] = s->conc[c] / (s->fracl[c] + p->kp * (1.0 - s->fracl[c]));
        }
        if (p->powderbed) {
            s->csfrac[c] = mx(s->csfrac[c], s->fracl[c]);
            s->solfrac[c] = mx(s->solfrac[c], s->fracl[c] * track);
            if (s->temp[c] > p->tsolid) s->mpphase[c] = (double)(int)trackid;
            if (s->temp[c] > p->tliquid) s->mpliqphase[c] = (double)(int)trackid;
        }
    }
}

/* driver step, src/main.cpp:125-141 */
void orc_begin_step(orc *s) {
    const amcfd_params *p = &s->p;
    if (p->n_dt_table > 0) {
        if (s->timet > s->dtt[3 * s->dt_state] && s->dt_state < p->n_dt_table - 1) s->dt_state += 1;
        s->p.delt = s->dtt[3 * s->dt_state + 1];
        s->p.outtime = s->dtt[3 * s->dt_state + 2];
    }
    s->timet += s->p.delt;
    if (!p->steady) update_laser(s);
    if (p->movingmesh) move_mesh(s);   /* main.cpp:144-145 */
    s->step++;
}

int orc_finished(const orc *s) { return !(s->timet < s->p.finaltime); }

int orc_run(orc *s, int max_steps) {
    int n = 0;
    while (!orc_finished(s) && (max_steps < 0 || n < max_steps)) {
        orc_begin_step(s);
        orc_solve(s);
        orc_end_step(s);
        ++n;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* set-up and access                                                         */
/* ------------------------------------------------------------------------ */
static double *dup_table(const double *src, int rows, int cols) {
    if (!src || rows <= 0) return NULL;
    double *d = (double *)malloc(sizeof(double) * rows * cols);
    memcpy(d, src, sizeof(double) * rows * cols);
    return d;
}

orc *orc_create(const amcfd_params *p) {
    orc *s = (orc *)calloc(1, sizeof(orc));
    s->p = *p;
    s->tool = dup_table(p->toolpath, p->n_toolpath, 6);
    s->dtt = dup_table(p->dt_table, p->n_dt_table, 3);
    s->heatt = dup_table(p->heat_table, p->n_heat_table, 6);
    s->ni = p->ni;
    s->nj = p->nj;
    s->nkf = p->nk;
    s->nim1 = s->ni - 1;
    s->nim2 = s->ni - 2;
    s->njm1 = s->nj - 1;
    s->njm2 = s->nj - 2;
    set_active_nk(s, s->nkf);
    size_t nn = (size_t)s->ni * s->nj * s->nkf;
    int n1 = s->ni + s->nj + s->nkf + 8;
    double **one[] = {&s->x, &s->y, &s->z, &s->xu, &s->yv, &s->zw, &s->dxpwinv, &s->dypsinv,
                      &s->dzpbinv, &s->fracx, &s->fracy, &s->fracz};
    for (size_t a = 0; a < sizeof(one) / sizeof(one[0]); ++a) *one[a] = zalloc(n1);
    double **three[] = {&s->volume, &s->volume_u, &s->volume_v, &s->volume_w, &s->vis, &s->diff,
                        &s->uvel, &s->vvel, &s->wvel, &s->enthalpy, &s->temp, &s->pp, &s->hnot,
                        &s->fracl, &s->fraclnot, &s->sourceinput, &s->csfrac, &s->rho, &s->solfrac,
                        &s->mpphase, &s->mpliqphase, &s->unot, &s->vnot, &s->wnot, &s->pressure,
                        &s->ap, &s->an, &s->as, &s->ae, &s->aw, &s->at, &s->ab, &s->apnot,
                        &s->dux, &s->dvy, &s->dwz, &s->su, &s->sp, &s->conc, &s->concnot, &s->concl,
                        &s->grad, &s->rate};
    for (size_t a = 0; a < sizeof(three) / sizeof(three[0]); ++a) *three[a] = zalloc(nn);
    s->areaij = zalloc((size_t)s->nj * s->ni);
    s->areauij = zalloc((size_t)s->nj * s->ni);
    s->areavij = zalloc((size_t)s->nj * s->ni);
    s->heatin = zalloc((size_t)s->nj * s->ni);
    s->areaik = zalloc((size_t)s->nkf * s->ni);
    s->areawik = zalloc((size_t)s->nkf * s->ni);
    s->areauik = zalloc((size_t)s->nkf * s->ni);
    s->areajk = zalloc((size_t)s->nkf * s->nj);
    s->fi = zalloc((size_t)s->nj * s->ni);
    s->zzm = zalloc((size_t)s->nj * s->ni);
    s->apdis = zalloc((size_t)s->nj * s->ni);
    s->zr = zalloc(nn);
    s->tempr = zalloc(nn);
    s->areavjk = zalloc((size_t)s->nkf * s->nj);
    s->areawjk = zalloc((size_t)s->nkf * s->nj);
    s->powerindicator = 1;
    s->laser_init = 1;
    s->trackindx = 1;
    s->scanvel = p->scanvel;
    generate_grid(s);
    initialize(s);
    /* FreeSurfaceManager::initializeFreeSurface, :35-67 (always run) */
    for (size_t c = 0; c < nn; ++c) {
        s->zr[c] = s->z[c / ((size_t)s->ni * s->nj)];
        s->tempr[c] = s->temp[c];
    }
    set_active_nk(s, p->nk_active0);
    if (p->steady) update_laser(s);
    return s;
}

void orc_destroy(orc *s) {
    if (!s) return;
    double *all[] = {s->x, s->y, s->z, s->xu, s->yv, s->zw, s->dxpwinv, s->dypsinv, s->dzpbinv, s->fracx,
                     s->fracy, s->fracz, s->volume, s->volume_u, s->volume_v, s->volume_w, s->vis, s->diff,
                     s->uvel, s->vvel, s->wvel, s->enthalpy, s->temp, s->pp, s->hnot, s->fracl, s->fraclnot,
                     s->sourceinput, s->csfrac, s->rho, s->solfrac, s->mpphase, s->mpliqphase, s->unot,
                     s->vnot, s->wnot, s->pressure, s->ap, s->an, s->as, s->ae, s->aw, s->at, s->ab,
                     s->apnot, s->dux, s->dvy, s->dwz, s->su, s->sp, s->areaij, s->areauij, s->areavij,
                     s->heatin, s->areaik, s->areawik, s->areauik, s->areajk, s->areavjk, s->areawjk,
                     s->tool, s->dtt, s->heatt, s->conc, s->concnot, s->concl, s->grad, s->rate,
                     s->fi, s->zzm, s->apdis, s->zr, s->tempr};
    for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); ++a) free(all[a]);
    free(s);
}

double *orc_field(orc *s, const char *n) {
    struct { const char *name; double *ptr; } t[] = {
        {"enthalpy", s->enthalpy}, {"temp", s->temp}, {"fracl", s->fracl}, {"uvel", s->uvel},
        {"vvel", s->vvel}, {"wvel", s->wvel}, {"pressure", s->pressure}, {"pp", s->pp},
        {"rho", s->rho}, {"diff", s->diff}, {"vis", s->vis}, {"hnot", s->hnot},
        {"fraclnot", s->fraclnot}, {"unot", s->unot}, {"vnot", s->vnot}, {"wnot", s->wnot},
        {"sourceinput", s->sourceinput}, {"dux", s->dux}, {"dvy", s->dvy}, {"dwz", s->dwz},
        {"csfrac", s->csfrac}, {"solfrac", s->solfrac}, {"mpPhase", s->mpphase},
        {"mpLiqPhase", s->mpliqphase}, {"ap", s->ap}, {"an", s->an}, {"as", s->as}, {"ae", s->ae},
        {"aw", s->aw}, {"at", s->at}, {"ab", s->ab}, {"su", s->su}, {"sp", s->sp},
        {"apnot", s->apnot}, {"heatin", s->heatin}, {"x", s->x}, {"y", s->y}, {"z", s->z},
        {"xu", s->xu}, {"yv", s->yv}, {"zw", s->zw}, {"volume", s->volume}, {"areaij", s->areaij},
        {"areajk", s->areajk}, {"areaik", s->areaik}, {"dxpwinv", s->dxpwinv},
        {"dypsinv", s->dypsinv}, {"dzpbinv", s->dzpbinv}, {"fracx", s->fracx}, {"fracy", s->fracy},
        {"fracz", s->fracz}, {"avgconcentration", s->conc}, {"avgconcentrationnot", s->concnot},
        {"concentration", s->concl}, {"grad", s->grad}, {"rate", s->rate}, {"zr", s->zr},
        {"tempr", s->tempr}, {"fi", s->fi}};
    for (size_t a = 0; a < sizeof(t) / sizeof(t[0]); ++a)
        if (strcmp(t[a].name, n) == 0) return t[a].ptr;
    return NULL;
}

void orc_get_scalars(const orc *s, amcfd_scalars *o) {
    memset(o, 0, sizeof(*o));
    o->timet = s->timet;
    o->niter = s->niter;
    o->itertot = s->itertot;
    o->resorh = s->resorh;
    o->resorm = s->resorm;
    o->resoru = s->resoru;
    o->resorv = s->resorv;
    o->resorw = s->resorw;
    o->tpeak = s->tpeak;
    o->alen = s->alen;
    o->depth = s->depth;
    o->width = s->width;
    o->ratio = s->ratio;
    o->heatinlaser = s->heatinlaser;
    o->heatout = s->heatout;
    o->accul = s->accul;
    o->heatvol = s->heatvol;
    o->ahtoploss = s->ahtoploss;
    o->beamposx = s->beamposx;
    o->beamposy = s->beamposy;
    o->scanvel = s->scanvel;
    o->powerindicator = s->powerindicator;
    o->trackindx = s->trackindx;
    o->istat = s->istat;
    o->iend = s->iend;
    o->jstat = s->jstat;
    o->jend = s->jend;
    o->kstat = s->kstat;
    o->istart = s->istart;
    o->jstart = s->jstart;
    o->nk = s->nk;
    o->step = s->step;
    o->delt = s->p.delt;
    o->outtime = s->p.outtime;
    o->resorc = s->resorc;
}

void orc_dims(const orc *s, int *ni, int *nj, int *nk) {
    *ni = s->ni;
    *nj = s->nj;
    *nk = s->nkf;
}
