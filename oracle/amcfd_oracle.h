/*
 * amcfd_oracle.h - CPU restatement of the reference's per-time-step SIMPLE
 * solve, in plain sequential C.  TEST INFRASTRUCTURE ONLY: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark use it as the checker;
 * the product library never links or calls it.
 *
 * It follows the reference's single-thread program order (the only
 * deterministic order of the reference, whose OpenMP sweeps race when
 * threaded) so that FP64 results agree with the reference to rounding.
 *
 * Parity pinning (what checks each function):
 *   every function on the path  - per-step full-field golden fixtures written
 *                                 by the compiled reference (tests/golden/):
 *                                 bare plate and powder bed, single track and
 *                                 toolpath, variable time steps, heat-source
 *                                 tables, steady mode, moving mesh (with the
 *                                 per-step grids), species and G/R, the
 *                                 energy-based free surface (with its monitor)
 *   grid                        - closed form (uniform zones, zone ends)
 *   H->T                        - inverse of h(T) = acpa/2 T^2 + acpb T
 *   power law                   - closed form, zero for |P| >= 10
 *   line TDMA                   - dense solve on small random lines
 *   block correction            - column residual sums vanish after it
 *   exponential scheme          - parity unpinned: pe/(exp(pe)-1) is 0/0 at
 *                                 pe = 0, so every deck using it is NaN in the
 *                                 reference and no fixture can pin it
 */
#ifndef AMCFD_ORACLE_H
#define AMCFD_ORACLE_H

#include "../include/amcfd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc  *orc_create(const amcfd_params *p);
void  orc_destroy(orc *s);
void  orc_begin_step(orc *s);      /* dt table, timet += dt, updateLaser  */
int   orc_solve(orc *s);           /* iterativeLoopSolve; 1 if maxit hit  */
void  orc_end_step(orc *s);        /* updateSolutions                     */
int   orc_run(orc *s, int max_steps);  /* whole time loop; returns steps */
int   orc_finished(const orc *s);
double *orc_field(orc *s, const char *name);   /* NULL if unknown name */
void  orc_get_scalars(const orc *s, amcfd_scalars *out);
void  orc_dims(const orc *s, int *ni, int *nj, int *nk);

/* unit-level pieces, exposed for closed-form tests */
double orc_disc_power(double pe);
double orc_disc_exponential(double pe);
double orc_h_to_t(const amcfd_params *p, double h, double *fl);
/* one x-line Thomas solve in the reference's form: a_w/a_p/a_e/d over
 * i=1..n, with x[0] and x[n+1] fixed; x has n+2 entries. */
void  orc_thomas(int n, const double *aw, const double *ap, const double *ae,
                 const double *d, double *x);
/* run only the energy half of one SIMPLE iteration (BCs .. melt pool) */
void  orc_energy_iteration(orc *s);
/* energy-based free surface at an output time (FreeSurfaceManager::execute);
 * returns 1 where the reference prints its bracket warning */
int   orc_free_surface(orc *s);
void  orc_free_surface_monitor(const orc *s, double *xmid, double *fmid, double *fimin,
                               double *fimax, double *vd, int *niter);

#ifdef __cplusplus
}
#endif
#endif
