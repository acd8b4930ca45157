// Internal declarations of the MI355X SIMPLE-solve library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/amcfd.h"

namespace amcfd {

int set_error(int code, const std::string &msg);

// ---------------------------------------------------------------------------
// parsed input deck (Mesh + DomainManager inputs)
// ---------------------------------------------------------------------------
struct Flags {
    int navierstokes = 1, steady = 0, isevap = 0, powderbed = 0, udtoolpath = 0, movingmesh = 0;
    int species = 0, solidification = 0, energyfreesurface = 0, outputsection = 0;
    int gaussheatsource = 0, volheatsource = 0, volheatsourcetable = 0;
};

struct Deck {
    amcfd_params p{};
    Flags flags;
    std::string keyword_id, mesh_file, tool_file, heat_file;
    // output settings (Mesh.cpp:263-268, 314-336, 407-411)
    std::string directory, geometry_out;
    std::vector<std::string> scalar_out, vector_out;
    int tecplot = 0;
    std::vector<std::vector<double>> tool_rows, dt_rows, heat_rows;
    std::vector<double> flat_tool, flat_dt, flat_heat;
    int load(const std::string &path);
    int read_mesh();
    void finalize_tables();
};

// ---------------------------------------------------------------------------
// device-side scalar block: everything the host reads once per iteration
// ---------------------------------------------------------------------------
struct DevScalars {
    double resorh, resorm, resoru, resorv, resorw, refmom, resorc;
    double tpeak, alen, depth, width;
    double ratio, heatout, accul, heatvol, ahtoploss, heatinlaser;
    double fluxwest, fluxeast, fluxbottom, fluxtop, fluxsouth, fluxnorth;
    int istat, iend, jstat, jend, kstat, imin, imax, jmin, jmax, kmin;
    int error;      // nonzero: a bounded wait in the line-sweep kernel expired
    // device-side control of the SIMPLE loop (captured graphs): the step's
    // iteration count, set once converged or at maxit, whether the last test
    // passed, and the flow kernels' skip flag (done, or nothing molten)
    int niter, done, converged, noflow;
    int pad;
};


// geometry on the device (1-D and 2-D arrays of the reference grid)
struct DevGeom {
    double *x, *y, *z, *xu, *yv, *zw, *dxpwinv, *dypsinv, *dzpbinv, *fracx, *fracy, *fracz;
    double *areaij, *areauij, *areavij;   // [nj*ni]
    double *areaik, *areawik, *areauik;   // [nk*ni]
    double *areajk, *areavjk, *areawjk;   // [nk*nj]
};

// solution + coefficient fields (ni*nj*nkf each)
struct DevFields {
    double *enthalpy, *temp, *fracl, *uvel, *vvel, *wvel, *pressure, *pp, *rho, *diff, *vis;
    double *hnot, *fraclnot, *unot, *vnot, *wnot, *sourceinput, *dux, *dvy, *dwz;
    double *csfrac, *solfrac, *mpphase, *mpliqphase;
    double *ap, *an, *as, *ae, *aw, *at, *ab, *su, *sp, *apnot;
    double *lpr, *lden, *ly;   // line-sweep pivots: pr, pivot, 1/pivot
    double *heatin;   // [nj*ni]
    // species (with *SOLVE_SPECIES) and G/R (with *CALCULATE_SOLIDIFICATION_PARAMETERS), else null
    double *conc, *concnot, *concl, *grad, *rate;
    // energy-based free surface (with *TRACK_ENERGY_BASED_FREE_SURFACE), else null:
    // surface shift and its previous iterate [nj*ni], node heights and temperatures
    double *fi, *zzm, *zr, *tempr;
};

// dimensions and constants every kernel needs (passed by value)
struct Dims {
    int ni, nj, nkf;              // allocation
    int nk, nim1, njm1, nkm1, nim2, njm2, nkm2;
    size_t sj, sk;                // strides in j and k
};

struct Box {
    int istat, iend, jstat, jend, kstat;
    int istatp1() const { return istat + 1; }
    int iendm1() const { return iend - 1; }
};

// per-iteration energy-source inputs (heat-source table entry, laser)
struct SourceParams {
    double avolpow, apowseta, heatthick, heatrb, avolfact;
    double beamposx, beamposy;
    int powerindicator;
};

// per-step inputs of a captured SIMPLE loop, written before its launches
struct StepDev {
    SourceParams sp;
    int powerindicator, maxit, istart, jstart;
};

// Gaussian surface flux of one laser position (BoundCondManager.cpp:128-131)
struct LaserParams {
    double beamx, beamy, alasfact, rb2, varlas, peakhin, pind;
};

// partial-sum buffers (deterministic two-pass reductions)
struct Partials {
    double *energy;     // [nblk_e][2]       resid sumd, sumh
    double *species;    // [nblk_e][2]       species resid sumd, sumc
    double *blk;        // [4][ni]           block-correction column sums
    double *h2t;        // [nblk_t][10]      tpeak + fluxes + accul/heatvol
    double *flow;       // [4][nblk_f][3]    per variable: sumd / umax / denom
    double *bc;         // [nblk_b]          ahtoploss
    double *laser;      // [nj*ni]           areaij * heatin, summed in order
    double *delphi;     // [ni]
    int nblk_e, nrow_e, nblk_t, nblk_f, nblk_b, nblk_l;
};

// mixed-precision mode: FP32 coefficients of the enthalpy equation, the
// block-correction partials and the second enthalpy buffer of the y/z passes
struct MxFields {
    float *ap, *an, *as, *ae, *aw, *at, *ab;
    double *su;        // FP64 source (the energy equation's f.su)
    double *blkpart;   // [4][ntile][ni] per-tile column partials
    double *blkfold;   // [4][split][ni]
    double *delphi;    // [ni]
    double *halt;      // second enthalpy buffer (y/z passes read one, write the other)
};

// line-sweep schedule (x-lines in Gauss-Seidel program order)
struct SweepPlan {
    int *tasks = nullptr;          // device: packed (pass, k, j) in wavefront order
    int ntasks = 0, width = 0, cap = 0;
    int kbeg = 0, kend = -1, jbeg = 0, jend = -1;   // cached shape
};

// ---------------------------------------------------------------------------
// kernel launchers (defined in the .hip files)
// ---------------------------------------------------------------------------
// with save, the ghost values the launch overwrites are kept there and *spec
// tells whether it ran (captured loops that run it ahead of the previous
// iteration's convergence test)
void launch_bc_enthalpy(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                        const amcfd_params &p, const int *skip = nullptr, double *save = nullptr,
                        int *spec = nullptr);
// the number of ghost values the enthalpy boundary launch sets
inline size_t bc_ghost_values(const Dims &d) {
    return (size_t)d.njm2 * d.nim2 + 2 * (size_t)d.nkm2 * d.njm2 + 2 * (size_t)d.nkm2 * d.nim2;
}
void launch_energy_assemble(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                            const amcfd_params &p, const SourceParams &sp);
void launch_laser_map(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                      const LaserParams &q, DevScalars *sc);
void launch_props(hipStream_t st, const Dims &d, DevFields f, const amcfd_params &p, const int *skip = nullptr);
// x-direction block correction of phi from the column-sum terms the assemble
// kernel left in f.sp / f.apnot (enthalpy or concentration)
void launch_block_correction(hipStream_t st, const Dims &d, DevFields f, Partials pr, double *phi, int shared);
// src_zero: the deck has no volumetric source, so the source field is zero
// and is not read (mixed mode)
void launch_h2t(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                const amcfd_params &p, const int *skip = nullptr, bool src_zero = false);
// launch_props then launch_h2t in one kernel (the same results), where nothing
// between them reads or writes what they read or write
void launch_props_h2t(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                      const amcfd_params &p, const int *skip = nullptr, bool src_zero = false);
// with device control (ctl), also sets sc->noflow for the flow kernels
void launch_melt_pool(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                      const amcfd_params &p, int istart, int jstart, DevScalars *sc, bool ctl = false,
                      const StepDev *step = nullptr);
// how a melt-pool-box kernel is launched: with dbox null, over the host box
// b with one workgroup per tile (the FP64 mode); otherwise over the box the
// melt-pool kernel left at dbox (istat, iend, jstat, jend, kstat), with
// nblocks workgroups striding over its tiles, skipped while *skip is set
struct BoxLaunch {
    const int *dbox = nullptr;
    const int *skip = nullptr;
    int nblocks = 0;
};
void launch_clean_velocities(hipStream_t st, const Dims &d, DevFields f, const Box &b, double tsolid,
                             const BoxLaunch &bl = BoxLaunch());
// the pass buffers of the persistent box passes (mixed mode) that a
// launch prepares: every box cell of passes buffers bstride apart to the
// pending pattern, the task counter to zero
struct BoxPrep {
    double *bufs = nullptr;
    size_t bstride = 0;
    int passes = 0;
    unsigned int *claim = nullptr;
};
// with prep (and a device box) the same launch also prepares the box
// passes' buffers
void launch_bc_flow(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, const Box &b,
                    const amcfd_params &p, int ivar, const BoxLaunch &bl = BoxLaunch(),
                    const BoxPrep *prep = nullptr);
void launch_momentum(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                     const Box &b, const amcfd_params &p, double scanvel, int ivar, const BoxLaunch &bl = BoxLaunch());
void launch_pressure(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                     const Box &b, const amcfd_params &p, const BoxLaunch &bl = BoxLaunch());
void launch_correct(hipStream_t st, const Dims &d, DevFields f, const Box &b, const amcfd_params &p,
                    const BoxLaunch &bl = BoxLaunch());
// with ctl and step, the same launch then applies the convergence test of
// CFDSolverManager.cpp:690-716 on the device: counts the iteration, sets
// done / converged
void launch_finalize(hipStream_t st, const Dims &d, Partials pr, DevScalars *sc, const amcfd_params &p,
                     int flow_done, int species_done, int steady, double width, bool ctl = false,
                     const StepDev *step = nullptr);
// grouped runs: out = whether all n flags done[i] are set
void launch_group_done(hipStream_t st, int *const *done, int n, int *out);
// resets the device loop control before a step's captured iterations
void launch_loop_reset(hipStream_t st, DevScalars *sc);
void launch_end_step(hipStream_t st, const Dims &d, DevFields f, const amcfd_params &p, double track,
                     double trackid);
// pivots of every line of a sweep and the pass buffers' "pending" fill (must
// precede launch_line_sweeps on the stream)
void launch_line_prep(hipStream_t st, const Dims &d, DevFields f, double *pb, size_t pstride, const SweepPlan &plan,
                      int ibeg, int iend, int guard);
// one Gauss-Seidel x-line sweep; pb holds four pass buffers of pstride
// doubles each (the field's size); scan selects the fast-mode line solve;
// shared selects the one-wave-per-line form (solvers sharing the device)
void launch_line_sweeps(hipStream_t st, const Dims &d, DevFields f, double *phi, double *pb, size_t pstride,
                        const SweepPlan &plan, int ibeg, int iend, int guard, int scan, unsigned long long *counter,
                        unsigned long long counter_base, DevScalars *sc, int nblocks, int shared);

// the device exp (fn 0) or log (fn 1) of n host values (tests); nonzero on a HIP failure
int eval_libm_device(int fn, const double *x, double *y, size_t n);

// mixed-precision mode (mixed.hip, energy.hip)
// split_mode with the device box split_box: 1 assembles the tiles away from
// the box (alongside the previous iteration's flow part), 2 the tiles near
// it, and when the loop has ended puts back the ghost values gsave holds if
// *gspec says the boundary launch ran ahead
void launch_energy_assemble_mixed(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                                  const MxFields &m, const amcfd_params &p, const SourceParams &sp, const int *skip,
                                  const StepDev *step = nullptr, const int *split_box = nullptr, int split_mode = 0,
                                  const double *gsave = nullptr, const int *gspec = nullptr, int split_blocks = 0);
int mixed_tile_count(const Dims &d);
// the mixed assembly writes the volumetric-source field only for decks that have one
int mixed_has_volumetric_source(const amcfd_params &p);
int mixed_max_line();
int mixed_max_yz();
size_t mixed_yz_lds(int len);   // LDS bytes of one y/z line workgroup
int mixed_blk_split();
void launch_mixed_block_correction(hipStream_t st, const Dims &d, const MxFields &m, double *phi, const int *skip);
// x-lines of one colour ((j + k) % 2 == colour), in place; line_batch
// consecutive lines per workgroup
void launch_mixed_xline(hipStream_t st, const Dims &d, const MxFields &m, double *phi, int ibeg, int iend, int jbeg,
                        int jend, int kbeg, int kend, int colour, const int *skip, int line_batch = 1);
// y-lines (dir 1) or z-lines (dir 2), hin -> hout
void launch_mixed_yzline(hipStream_t st, const Dims &d, const MxFields &m, const double *hin, double *hout, int dir,
                         int ibeg, int iend, int jbeg, int jend, int kbeg, int kend, const int *skip);
// melt-pool box x-lines of one colour (FP64 coefficients, reference guard)
// all red-black passes of one melt-pool box variable in one persistent
// launch (plus, with prep, the launch that prepares its pass buffers;
// without, an earlier launch has prepared them); nonzero if a launch failed
int launch_box_passes(hipStream_t st, const Dims &d, DevFields f, double *phi, double *bufs, size_t bstride,
                      int passes, unsigned int *claim, int *error, const int *skip, const int *dbox, int nblocks,
                      bool prep = true);
void launch_box_xline(hipStream_t st, const Dims &d, DevFields f, double *phi, int ibeg, int iend, int jbeg, int jend,
                      int kbeg, int kend, int colour, const int *skip, const int *dbox = nullptr, int nblocks = 0);

// species (species.hip)
void launch_bc_species(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, const amcfd_params &p);
void launch_species_assemble(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, Partials pr,
                             const amcfd_params &p);
void launch_species_end(hipStream_t st, const Dims &d, DevFields f, const amcfd_params &p);
void launch_solidification(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, double tsolid,
                           double scanvel);

// energy-based free surface at an output time (surface.hip); vd = added
// droplet volume; out[6] receives multiplier, residual, min/max shift,
// bisection count and the bracket-warning flag
void launch_free_surface(hipStream_t st, const Dims &d, const DevGeom &g, DevFields f, const amcfd_params &p,
                         double vd, double *out);

// moving mesh: remap the twelve (fourteen with species) solution fields from the node coordinates
// xold/yold onto g.x/g.y (x pass, then y pass), dirx/diry = sign of the move
void launch_remap(hipStream_t st, const Dims &d, DevFields f, const DevGeom &g, const double *xold,
                  const double *yold, int dirx, int diry, int species);

// wavefront task list for the Gauss-Seidel line sweeps (host side)
void build_sweep_tasks(int kbeg, int kend, int jbeg, int jend, std::vector<int> &tasks, int &width);

}  // namespace amcfd
