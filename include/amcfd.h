/*
 * amcfd.h - C ABI of the MI355X SIMPLE solve (AM-CFD per-time-step path).
 *
 * The reference exposes this path only as C++ member functions on heap
 * objects created by its executable:
 *   Mesh::getDomainInfo / readMeshFile        src/sources/Mesh.cpp:39-550
 *   DomainManager::assignInputValues          src/sources/DomainManager.cpp:29-263
 *   CFDSolverManager::generateGrid            src/sources/CFDSolverManager.cpp:34-302
 *   CFDSolverManager::initializeSolver        src/sources/CFDSolverManager.cpp:402-590
 *   BoundCondManager::updateLaser             src/sources/BoundCondManager.cpp:68-170
 *   CFDSolverManager::iterativeLoopSolve      src/sources/CFDSolverManager.cpp:595-722
 *   CFDSolverManager::updateSolutions         src/sources/CFDSolverManager.cpp:2625-2687
 *   SolutionVariableManager name->field map   src/headers/SolutionVariableManager.h:18-19
 * called in that order by src/main.cpp:56-162.  Each entry point below names
 * the member function it replaces.
 *
 * Conventions
 *   - status: 0 ok; > 0 informational bits; < 0 errors (AMCFD_E_*).
 *   - fields use the reference layout: index (k*nj + j)*ni + i, i fastest,
 *     FP64, ni*nj*nk values (nk = allocated node count in z).
 *   - caller owns host buffers, the library owns device memory.
 *   - one handle per host thread; calls are stream ordered and synchronous
 *     where they return host data.
 */
#ifndef AMCFD_H
#define AMCFD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMCFD_MAX_ZONES 32

/* status codes */
#define AMCFD_OK              0
#define AMCFD_I_MAXIT         1    /* solve ended at maxit without converging  */
#define AMCFD_E_ARG          -1    /* bad argument                             */
#define AMCFD_E_IO           -2    /* deck / mesh / table file not readable    */
#define AMCFD_E_PARSE        -3    /* deck grammar error                       */
#define AMCFD_E_UNSUPPORTED  -4    /* keyword outside this path                */
#define AMCFD_E_HIP          -5    /* HIP runtime failure                      */
#define AMCFD_E_NAN          -6    /* non-finite value detected on device      */
#define AMCFD_E_CAPACITY     -7    /* problem too large for the device         */
#define AMCFD_E_NAME         -8    /* unknown field name                       */

/* discretisation schemes, Mesh.cpp:378-392 (default upwind, Mesh.h:59) */
#define AMCFD_DISC_UPWIND      0
#define AMCFD_DISC_POWERLAW    1
#define AMCFD_DISC_EXPONENTIAL 2

/* solver modes */
#define AMCFD_MODE_FP64        0   /* plain FP64, reference program order     */
#define AMCFD_MODE_SCAN        1   /* FP64; x-line solves by parallel scans in
                                      the same Gauss-Seidel order: rounding
                                      differs from the reference's, so results
                                      agree to the iteration's sensitivity,
                                      not bit for bit                         */
#define AMCFD_MODE_MIXED       2   /* enthalpy coefficients stored in FP32;
                                      every line pass a delta-form correction
                                      (FP64 residual, FP32 line solve, FP64
                                      update) in parallel order: x-lines
                                      red-black, y- and z-lines alternating;
                                      H->T, heat balance, mass imbalance and
                                      the convergence test stay FP64.  Agrees
                                      with the FP64 mode to the tolerance of
                                      the converged iteration, not bit for bit */

/*
 * Every value the path reads from the reference's Mesh and DomainManager
 * members (src/headers/Mesh.h:41-68, src/headers/DomainManager.h:26-72).
 * Tables are borrowed pointers owned by whoever filled the struct.
 */
typedef struct amcfd_params {
    /* grid zones: DomainManager.cpp:33-46 */
    int    nzx, nzy, nzz;
    double xzone[AMCFD_MAX_ZONES], ncvx[AMCFD_MAX_ZONES], powrx[AMCFD_MAX_ZONES];
    double yzone[AMCFD_MAX_ZONES], ncvy[AMCFD_MAX_ZONES], powry[AMCFD_MAX_ZONES];
    double zzone[AMCFD_MAX_ZONES], ncvz[AMCFD_MAX_ZONES], powrz[AMCFD_MAX_ZONES];
    /* control: Mesh.cpp:284-299 */
    double finaltime, maxit, nonlintol, delt, outtime;
    /* *PARAMETERS: Mesh.cpp:440-470 */
    double grav, sigm, tempamb, pressamb, gasconst;
    /* feature flags */
    int    navierstokes;       /* 0 if *FLUID_FLOW_OFF                     */
    int    discretization;     /* AMCFD_DISC_*                             */
    int    steady;             /* delt >= 100, DomainManager.cpp:203-210   */
    int    isevap, powderbed, udtoolpath, gaussheatsource;
    int    volheatsource, volheatsourcetable;
    /* single track: DomainManager.cpp:51-54 */
    double lasertime, xstart, ystart, scanvel;
    /* material: DomainManager.cpp:59-73 */
    double dens, denl, viscos, tsolid, tliquid, hsmelt, hlfriz, acpa, acpb, acpl;
    double thconsa, thconsb, thconsc, thconla, thconlb;
    /* relaxation: DomainManager.cpp:76-80 */
    double urfu, urfv, urfw, urfp, urfh;
    /* energy / temperature / momentum BCs: DomainManager.cpp:83-100 */
    double htci, htcj, htck1, htckn, emiss;
    double tempwest, tempeast, tempnorth, tempbottom, temppreheat;
    double beta, dgdtp, darcyKo, a_oxygen;
    /* Gaussian surface laser: DomainManager.cpp:103-110 */
    double alaspow, alaseta, alasrb, alasfact;
    /* evaporation: DomainManager.cpp:113-120 */
    double evaplatent, evaptemp, evapfact, molarmass;
    /* volumetric source: DomainManager.cpp:165-186 */
    double avolpow, apowseta, heatthick, heatrb, avolfact;
    /* powder bed: DomainManager.cpp:147-162 (zzone/ncvz already extended) */
    int    numlayers, ncvzlayer;
    double layerheight, powderden, powdercpa, powdercpb, powderthcona, powderthconb;
    /* tables (row-major) */
    int    n_toolpath;   const double *toolpath;    /* t x y z on track, 6/row */
    int    n_dt_table;   const double *dt_table;    /* t_until dt outtime, 3/row */
    int    n_heat_table; const double *heat_table;  /* t P eta thick rb fact, 6/row */
    /* derived: DomainManager.cpp:213-254 */
    int    ni, nj, nk;         /* node counts incl. boundary nodes (nk = full) */
    int    nk_active0;         /* active nk after powder-layer deactivation    */
    double hlcal, hlatnt;
    /* *ACTIVATE_MOVING_MESH: Mesh.cpp:364-369 (0-based zone indices); the
     * zones either side of them stretch and shrink with the beam every step
     * (CFDSolverManager.cpp:72-124) */
    int    movingmesh, xmoveindex, ymoveindex;
    /* *SOLVE_SPECIES: DomainManager.cpp:123-130 (initial concentration, mass
     * diffusivity, partition coefficient, under-relaxation factor) */
    int    species;
    double initconcentration, massdiff, kp, urfc;
    /* *CALCULATE_SOLIDIFICATION_PARAMETERS: temperature gradient G and growth
     * rate R after every solve (CFDSolverManager.cpp:2763-2804) */
    int    solidification;
    /* *TRACK_ENERGY_BASED_FREE_SURFACE: DomainManager.cpp:133-144 (bisection
     * bracket, surface tension, energy-iteration limit and tolerance,
     * bisection limit and tolerance, relaxation); evaluated at output times */
    int    energyfreesurface;
    double initlam1, initlam2, surftens, maxiterenergy, tolenergy, maxiterbisect, tolbisect, urfree;
    /* *OUTPUT_SECTION: DomainManager.cpp:189-198 (box min, box max, interval) */
    int    outputsection;
    double gxmin, gymin, gzmin, gxmax, gymax, gzmax, toutsect;
} amcfd_params;

/* scalars the reference keeps as members (CFDSolverManager.h:81-83,
 * BoundCondManager.h:62-80) after a solve / step */
typedef struct amcfd_scalars {
    double timet;
    int    niter, itertot;
    double resorh, resorm, resoru, resorv, resorw;
    double tpeak, alen, depth, width;
    double ratio, heatinlaser, heatout, accul, heatvol, ahtoploss;
    double beamposx, beamposy, scanvel;
    int    powerindicator, trackindx;
    int    istat, iend, jstat, jend, kstat, istart, jstart, nk;
    int    step;
    double delt, outtime;   /* time step and output interval in force (main.cpp:125-133) */
    double resorc;          /* species residual (CFDSolverManager.cpp:2128); 0 without species */
} amcfd_scalars;

typedef struct amcfd_deck amcfd_deck;      /* parsed input deck          */
typedef struct amcfd_solver amcfd_solver;  /* device-resident solver     */

/* Parse a keyword deck and its mesh / toolpath / table files.
 * Replaces Mesh::Mesh + getDomainInfo + assignParameters + readMeshFile
 * (Mesh.cpp:23-26,39-550) and DomainManager::assignInputValues /
 * getToolpath / getHeatSourceParameters (DomainManager.cpp:29-312).
 * Relative file names resolve against the current directory, as in the
 * reference. */
int  amcfd_deck_load(const char *deck_path, amcfd_deck **out);
const amcfd_params *amcfd_deck_params(const amcfd_deck *deck);
void amcfd_deck_free(amcfd_deck *deck);

/* Create the device solver: generateGrid + initializeSolver +
 * initializeBoundaries (+ powder layer deactivation), PrintManager.cpp:49-74,
 * and the steady-state laser set-up of main.cpp:105-106. */
int  amcfd_create(const amcfd_params *params, int mode, amcfd_solver **out);
void amcfd_destroy(amcfd_solver *s);

/* One time step, split like src/main.cpp:125-162:
 *   begin_step = variable-dt table + timet += delt + updateLaser
 *   solve_step = CFDSolverManager::iterativeLoopSolve
 *   end_step   = CFDSolverManager::updateSolutions (transient only)
 * solve_step returns AMCFD_I_MAXIT when maxit was reached. */
int  amcfd_begin_step(amcfd_solver *s);
int  amcfd_solve_step(amcfd_solver *s);
int  amcfd_end_step(amcfd_solver *s);
/* begin + solve + end; returns 1 once timet >= finaltime (main.cpp:123) */
int  amcfd_step(amcfd_solver *s, int *finished);

/* Field access by registry name (CFDSolverManager.cpp:289-302,528-560):
 * enthalpy temp fracl uvel vvel wvel pressure pp rho diff vis hnot fraclnot
 * unot vnot wnot sourceinput dux dvy dwz csfrac solfrac.  n must equal
 * ni*nj*nk (full z extent). */
int  amcfd_get_field(amcfd_solver *s, const char *name, double *host, size_t n);
int  amcfd_set_field(amcfd_solver *s, const char *name, const double *host, size_t n);
/* 1-D / 2-D geometry: x y z xu yv zw dxpwinv dypsinv dzpbinv fracx fracy fracz
 * volume areaij areajk areaik */
int  amcfd_get_geometry(amcfd_solver *s, const char *name, double *host, size_t n);
int  amcfd_get_scalars(amcfd_solver *s, amcfd_scalars *out);
int  amcfd_dims(amcfd_solver *s, int *ni, int *nj, int *nk);

/* Device time of the last solve_step in milliseconds, split into the line
 * sweeps (x-line TDMA) and everything else; -1 if unavailable.
 * No reference counterpart: the reference prints only wall-clock totals
 * (PrintManager.cpp:713-737). */
int  amcfd_last_timing(amcfd_solver *s, double *total_ms, double *tdma_ms);
/* Record HIP events around the kernel families ("sweep_h", "sweep_u",
 * "sweep_v", "sweep_w", "sweep_p", "assemble_h", "h2t") and report the
 * accumulated launches and device milliseconds: on = 1 times every family,
 * 2 the whole-domain enthalpy sweep ("sweep_h") alone, 0 none.  No reference
 * counterpart. */
int  amcfd_set_profiling(amcfd_solver *s, int on);
int  amcfd_kernel_stats(amcfd_solver *s, const char *kernel, long *launches, double *total_ms);

/* Output files in the reference's formats, so a run leaves the files the
 * reference executable writes (src/main.cpp:115-241):
 *   write_vts       one StructuredGrid file, vtsBinWriter::visualOut
 *                   (vtsBinWriter.cpp:50-139); scalar_names are the deck's
 *                   *SCALAR_OUT names (uvel/vvel/wvel written cell-centred as
 *                   auvel/avvel/awvel; names the reference does not hold are
 *                   skipped, as it does, PrintManager.cpp:314-378)
 *   write_pvd       the ParaView collection of all files so far
 *                   (vtsBinWriter.cpp:420-468)
 *   append_residuals / append_final_geometry
 *                   PrintManager::outputResiduals (PrintManager.cpp:380-533)
 *                   and ::outputFinalGeometry (:658-677), appending to path */
int  amcfd_write_vts(amcfd_solver *s, const char *path, const char *const *scalar_names, int n_names);
int  amcfd_write_pvd(const char *path, int n, const double *times, const char *const *files);
int  amcfd_append_residuals(amcfd_solver *s, const char *path);
int  amcfd_append_final_geometry(amcfd_solver *s, const char *path);
/* Deck strings the output layer needs (Mesh.cpp:263-268, 314-336, 407-411):
 * key "keyword_id", "directory", "final_geometry", "mesh_file", or
 * "scalar_out" / "vector_out" with an index; NULL when absent. */
const char *amcfd_deck_string(const amcfd_deck *deck, const char *key, int index);

/* Energy-based free surface at an output time: FreeSurfaceManager::execute
 * (FreeSurfaceManager.cpp:72-220), called by main.cpp:175-176 and :211-212.
 * Updates the registry fields "zr" (shifted node heights, then used as the
 * VTS point heights) and "tempr"; appends the monitor block to monitor_path
 * when it is not NULL.  Needs *TRACK_ENERGY_BASED_FREE_SURFACE. */
int  amcfd_free_surface(amcfd_solver *s, const char *monitor_path);
/* Extracted section: PrintManager::outputSection (PrintManager.cpp:538-624),
 * called every step after the solve (main.cpp:155-156) and once with finish
 * set at the end (:214-215).  Accumulates the time step and, when the
 * *OUTPUT_SECTION interval has passed (or finish), appends the box's nodes
 * to <base>_ExtractedSection<layer>.txt; *written tells which. */
int  amcfd_output_section(amcfd_solver *s, const char *base, int finish, int *written);

/* Independent instances on the one GPU at once (a parameter sweep): every
 * solver advances up to max_steps time steps (to its final time if
 * max_steps < 0).  `workers` host threads (16 if <= 0, at most n) take the
 * instances largest grid first, each running one instance at a time on that
 * instance's own stream, so the instances' kernels and line sweeps overlap
 * on the device; every instance computes exactly what it computes alone.
 * steps_done[i] receives the steps instance i took.  On an error the first
 * failing instance's status is returned.  The device runs as many streams
 * side by side as the HIP runtime has hardware queues (GPU_MAX_HW_QUEUES, 4
 * by default).  No reference counterpart: the reference runs one deck per
 * process. */
int  amcfd_run_batch(amcfd_solver *const *s, int n, int max_steps, int workers, int *steps_done);

/* n mixed-mode solvers (captured loop on, x-line schedule, no heat-source
 * table) advanced in lock step, up to max_steps steps each (negative: until
 * each finishes): every kernel of a SIMPLE iteration is one launch over all
 * of them, the instances' tiles and line batches side by side in one grid,
 * captured with graph_block iterations into one graph that is relaunched
 * until every instance's convergence test (CFDSolverManager.cpp:690-716,
 * on the device) has ended its step.  Instances may differ in grid, deck
 * and parameters; each computes exactly what it computes alone.
 * steps_done[i] receives the steps instance i took.  Returns
 * AMCFD_E_UNSUPPORTED (before any step) for a solver that cannot take part.
 * No reference counterpart: the reference runs one deck per process. */
int  amcfd_run_grouped(amcfd_solver *const *s, int n, int max_steps, int *steps_done);

/* The reference's x-line TDMA over many independent lines held in device
 * memory: the recurrence of CFDSolverManager::solveEntireDomain
 * (CFDSolverManager.cpp:2500-2522; guard = 0) or, with the |den| <= 1e-12
 * guard, CFDSolverManager::solveLiquidDomain (:2421-2450), with the
 * reference's roundings.  Line l occupies entries [l*n, (l+1)*n) of every
 * array; phi[l*n] and phi[l*n+n-1] are the fixed end values, rows 1..n-2 are
 * solved from aw, ap, ae and the right-hand side d (su plus the neighbour
 * lines' terms).  Device pointers; stream is a hipStream_t (NULL: the null
 * stream); returns once the work is queued.  The pr/qr scratch (2*n*nlines
 * doubles) is a per-device workspace that the library allocates on first use
 * and grows; calls that overlap on one device share it and race. */
int  amcfd_tdma_lines(const double *aw, const double *ap, const double *ae, const double *d, double *phi,
                      int n, int nlines, int guard, void *stream);

/* amcfd_tdma_lines without any allocation, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  Same arguments
 * and results as amcfd_tdma_lines; the scratch is one fixed device array of
 * 2^21 doubles (16 MiB) per GPU, held until the process ends, so
 * n * nlines <= 2^20 (larger batches return AMCFD_E_CAPACITY before the
 * device is touched).  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU race. */
int  amcfd_tdma_lines_noalloc(const double *aw, const double *ap, const double *ae, const double *d,
                              double *phi, int n, int nlines, int guard, void *stream);

/* Mixed mode only: the order of the enthalpy line passes of one SIMPLE
 * iteration ('x' = x-lines red then black, in place; 'y' / 'z' = y- or
 * z-lines, one lane per line, between the field and a second buffer; an
 * even number of 'y'/'z'; default "xxxx") and the red-black x passes per
 * melt-pool box variable (default 4).  The reference sweeps x-lines only,
 * four Gauss-Seidel passes (CFDSolverManager.cpp:2415-2453, 2492-2525). */
int  amcfd_set_mixed_schedule(amcfd_solver *s, const char *schedule, int box_passes);

/* Mixed mode only: run each step's SIMPLE loop from a captured HIP graph of
 * `block` iterations (on = 1, the default) or from host launches (0).  In
 * the graph every kernel reads the loop's state from device memory: the
 * melt-pool box and the flow/no-flow decision (CFDSolverManager.cpp:657-
 * 675), and the convergence test (:690-716), which a device kernel takes
 * after each iteration and which turns the rest of the block into no-ops;
 * the host relaunches the graph until the step is done or maxit, reading
 * the scalars once per block.  box_blocks workgroups stride over the box
 * (default 256, at most 4096).  Both ways run the same launches, so the results agree bit
 * for bit.  A step with a heat-source table row change always runs from the
 * host. */
int  amcfd_set_graph(amcfd_solver *s, int on, int block, int box_blocks);
/* Mixed mode: the box passes of each melt-pool box variable (u, v, w, p')
 * in one persistent launch (on = 1, the default) or one launch per red or
 * black half pass (on = 0).  Alone the two take the same time; beside other
 * solvers (amcfd_run_batch) the persistent form's fewer launches measured
 * 30% more throughput.  The persistent form claims its line tasks in
 * pass order and hands line values over through per-pass buffers; its waits
 * are bounded and an expired one ends the step with AMCFD_E_HIP.  Both forms
 * give the same bits.  No reference counterpart. */
int  amcfd_set_box_fused(amcfd_solver *s, int on);
/* Mixed mode: consecutive x-lines each workgroup solves in the enthalpy's
 * red-black passes (1..64, default 1).  The lines of a half pass are
 * independent, so this changes the passes' speed and never their results.
 * No reference counterpart. */
int  amcfd_set_line_batch(amcfd_solver *s, int lines);
/* Mixed mode, captured loop: while one iteration's flow part (the melt-pool
 * box solves, latency-bound and on few compute units) runs, the next
 * iteration's enthalpy boundary launch and the tiles of its assembly away
 * from the box run beside it on a second stream (on = 1, with about 1.5
 * workgroups per compute unit striding over those tiles; on > 1 sets that
 * workgroup count); the tiles near the box follow the flow part.  A solver
 * starts from the tuning table's choice (off unless a row turns it on).  The
 * boundary launch keeps the ghost values it overwrites and they are put back
 * if the convergence test ends the loop first.  on = 0 runs everything in
 * order, as amcfd_run_batch does with several workers (there a stream that
 * waits on another holds up the hardware queue it shares with other
 * solvers).  Every setting gives the same bits.  No reference counterpart. */
int  amcfd_set_assembly_overlap(amcfd_solver *s, int on);
/* Mixed mode: the captured loop's current settings, the line batch and
 * whether the assembly overlap is on.  A mixed-mode solver starts from the
 * tuning table's row for its interior cell count (below). */
int  amcfd_get_graph(const amcfd_solver *s, int *on, int *block, int *box_blocks, int *box_fused, int *line_batch,
                     int *overlap);

/* The launch-shape table of the mixed mode (no reference counterpart),
 * written by am-cfd_amd/autotune.py: lines "max_cells graph_block
 * box_blocks [box_fused [line_batch [overlap]]]", '#' comments; the first
 * row with max_cells >= the grid's interior cells applies, else the built-in
 * 8, 256, 1, 1 and the assembly overlap (amcfd_set_assembly_overlap) off;
 * the shipped table turns the overlap on where it measured faster.  These
 * values change the captured loop's speed, never its results.  The
 * library's table is the file named by $AMCFD_TUNING, else tuning.txt beside
 * libamcfd.so, read once.
 * amcfd_tuning_choice reports the choice for `cells` from the file at
 * `path`, or from the library's table when path is NULL; it needs no GPU. */
int  amcfd_tuning_choice(const char *path, long long cells, int *graph_block, int *box_blocks, int *box_fused,
                         int *line_batch, int *overlap);

/* Copy the whole state of src into dst (same deck, any modes): every
 * solution field, the grid, the laser and time-loop state and the scalars,
 * so that dst continues exactly where src stands, e.g. to run the same
 * state on in another mode.  No reference counterpart (the reference has
 * one solver per process). */
int  amcfd_copy_state(amcfd_solver *dst, const amcfd_solver *src);

/* Test hook, no reference counterpart: raise the device error flag that
 * the Gauss-Seidel line sweeps' bounded waits watch, so that every wait of
 * the next sweep gives up at once, every kernel still ends, and the solve
 * returns AMCFD_E_HIP.  The flag stays raised for the solver's life. */
int  amcfd_test_expire_waits(amcfd_solver *s);

/* Test hook, no reference counterpart: evaluate the device's exp (fn 0) or
 * log (fn 1), the functions the kernels use where the reference calls the C
 * library's (std::exp at BoundCondManager.cpp:150, :604, :729 and
 * CFDSolverManager.cpp:1735, :3100; std::log at BoundCondManager.cpp:730),
 * on n host values x into y, so that tests can hold them to the C library
 * bit for bit. */
int  amcfd_test_libm(int fn, const double *x, double *y, size_t n);

/* Learned warm start (no reference counterpart): a 3-D convolutional
 * encoder-decoder (4 levels, 16-128 channels, 3x3x3 kernels, BF16 on the
 * matrix cores, FP32 accumulation) that predicts a step's temperature
 * increment and pressure change in a 32 x 32 x 16-cell window around the
 * beam from the state the step starts from and the step's laser flux.  It
 * is trained on the GPU from steps the solver runs, and then supplies the
 * step's initial iterate; the SIMPLE iteration still runs to the reference's
 * convergence test (CFDSolverManager.cpp:690-716), so only the number of
 * iterations changes.
 *   ws_create / destroy  network with seeded initial weights
 *   ws_attach            mode 0 off, 1 record one training sample per step
 *                        (the state before and the increment after the
 *                        solve), 2 start a step from the guess when the
 *                        step before it converged (a step that ends at
 *                        maxit depends on its start; runs whose steps end
 *                        there keep the plain start), 3 start every step
 *                        from the guess; one
 *                        network may serve many solvers, also solvers that
 *                        step on other host threads (amcfd_run_batch): each
 *                        keeps its own pending sample, and the network's
 *                        forward passes run one at a time.  Destroying the
 *                        network detaches its solvers; ws_train and ws_load
 *                        must not overlap steps of attached solvers
 *   ws_train             Adam over random batches of the recorded samples;
 *                        loss of the first and last batch
 *   ws_set_factors       fractions of the predicted temperature and pressure
 *                        changes applied (default 1 and 0)
 *   ws_save / ws_load    FP32 weights
 *   ws_param_count / ws_params (set = 0 read, 1 write) / ws_loss_grad
 *                        the flat parameter vector (per layer: weights
 *                        [co][tap][ci], then biases) and the loss and its
 *                        gradient on host tensors, NDHWC with 16 input
 *                        channels and 2 target channels (tests) */
typedef struct amcfd_ws amcfd_ws;
int  amcfd_ws_create(int seed, amcfd_ws **out);
void amcfd_ws_destroy(amcfd_ws *ws);
int  amcfd_ws_attach(amcfd_solver *s, amcfd_ws *ws, int mode);
/* steps of this solver started from the guess, and steps mode 2 kept on the
 * plain start because the step before did not converge */
int  amcfd_ws_counts(const amcfd_solver *s, int *applied, int *guarded);
int  amcfd_ws_train(amcfd_ws *ws, int iters, int batch, double lr, double *loss_first, double *loss_last);
int  amcfd_ws_samples(amcfd_ws *ws);
int  amcfd_ws_set_factors(amcfd_ws *ws, double tfactor, double pfactor);
int  amcfd_ws_save(amcfd_ws *ws, const char *path);
int  amcfd_ws_load(amcfd_ws *ws, const char *path);
int  amcfd_ws_param_count(amcfd_ws *ws);
int  amcfd_ws_params(amcfd_ws *ws, float *host, size_t n, int set);
int  amcfd_ws_loss_grad(amcfd_ws *ws, const float *x, const float *target, const float *mask, int n, int d, int h,
                        int w, double *loss, float *grad);

/* Thread-local description of the last error. */
const char *amcfd_last_error(void);
const char *amcfd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* AMCFD_H */
