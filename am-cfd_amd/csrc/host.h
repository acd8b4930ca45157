// Host-side solver object: owns the device state and drives one time step
// (laser update, SIMPLE iterations, push-back) through the HIP kernels.
#pragma once

#include "group.h"
#include "internal.h"

#include <map>
#include <memory>

namespace amcfd {

struct HostGrid {
    int ni = 0, nj = 0, nk = 0;
    std::vector<double> x, y, z, xu, yv, zw, dxpwinv, dypsinv, dzpbinv, fracx, fracy, fracz;
    std::vector<double> areaij, areauij, areavij, areaik, areawik, areauik, areajk, areavjk, areawjk;
    void build(const amcfd_params &p);
    // CFDSolverManager::generateGrid with the z extent nka (the active nk)
    void regenerate(const amcfd_params &p, int nka);
    double volume(int i, int j, int k) const;
};

struct InitialState {
    std::vector<double> enthalpy, hnot, temp, rho, vis, diff, csfrac;
};

void initial_fields(const amcfd_params &p, const HostGrid &g, InitialState &s);

int energy_partial_sizes(const Dims &d, int *nblk_e, int *nrow_e, int *nblk_t);
int flow_partial_blocks(const Box &b, const Dims &d);
int sweep_max_line();

class WarmStart;
class Solver;
int ws_capture_input(WarmStart *w, Solver &s);
int ws_capture_target(WarmStart *w, Solver &s);
int ws_apply(WarmStart *w, Solver &s);
void ws_attach(WarmStart *w, Solver *s);
// the tuning table's choice for a grid of `cells` interior cells (tuning.cpp);
// leaves the values alone when no table or no row applies
void tuning_lookup(long long cells, int *graph_block, int *box_blocks, int *box_fused, int *line_batch,
                   int *overlap);
void ws_detach(WarmStart *w, Solver *s);

// accumulated device time of one kernel family (HIP events on the solver stream)
struct KernelTimer {
    long launches = 0;
    double total_ms = 0.0;
};

class Solver {
  public:
    ~Solver();
    int create(const amcfd_params &p, int mode);
    int begin_step();
    int solve_step();
    int end_step();
    int finished() const { return !(timet_ < p_.finaltime); }
    int get_field(const char *name, double *host, size_t n);
    int set_field(const char *name, const double *host, size_t n);
    int get_geometry(const char *name, double *host, size_t n);
    int get_scalars(amcfd_scalars *out);
    void dims(int *ni, int *nj, int *nk) const { *ni = d_.ni; *nj = d_.nj; *nk = d_.nkf; }
    // read-only views for the output writers (output.cpp)
    const DevScalars &host_scalars() const { return *sc_host_; }
    const amcfd_params &params() const { return p_; }
    const HostGrid &grid() const { return hg_; }
    int active_nk() const { return d_.nk; }
    double last_total_ms = -1.0, last_tdma_ms = -1.0;
    void set_profiling(int on) { profile_ = on; }
    // the solver runs beside others on the device (amcfd_run_batch)
    void set_shared_device(bool on) { shared_device_ = on; }
    int mode() const { return mode_; }
    void graph_settings(int *on, int *block, int *box_blocks) const {
        *on = use_graph_;
        *block = graph_block_;
        *box_blocks = box_blocks_;
    }
    int box_fused() const { return box_fused_; }
    int set_box_fused(int on);
    // mixed mode: consecutive x-lines per workgroup in the enthalpy's passes
    int line_batch() const { return line_batch_; }
    int set_line_batch(int lines);
    int set_assembly_overlap(int on);
    int assembly_overlap() const { return fork_asm_ ? 1 : 0; }
    // mixed mode: order of the enthalpy line passes ('x' red-black x-lines,
    // 'y' y-lines, 'z' z-lines; an even number of y/z) and red-black x
    // passes per melt-pool box variable
    int set_mixed_schedule(const char *sched, int box_passes);
    // mixed mode: run the SIMPLE loop from a captured graph of `block`
    // iterations (on, the default) or from host launches (off); box_blocks
    // workgroups for the melt-pool-box kernels
    int set_graph(int on, int block, int box_blocks);
    // raise the sweeps' device error flag (test hook for the bounded waits)
    int raise_sweep_error();
    // take over another solver's state (same deck): every solution field on
    // the device, the grid, the laser / time-loop state and the scalars
    int copy_state_from(const Solver &o);
    // learned warm start: 0 off, 1 record a training sample per step, 2 use
    // the network's prediction as the initial iterate of each step whose
    // predecessor converged, 3 on every step
    // the network keeps per-solver recording state: attaching registers this
    // solver with it, detaching (or destroying the solver) releases that
    void attach_warm_start(WarmStart *w, int mode) {
        WarmStart *nw = mode ? w : nullptr;
        if (ws_ && ws_ != nw) ws_detach(ws_, this);
        if (nw && nw != ws_) ws_attach(nw, this);
        ws_ = nw;
        ws_mode_ = mode;
    }
    void warm_start_counts(int *applied, int *guarded) const {
        *applied = ws_applied_;
        *guarded = ws_guarded_;
    }
    // called by a network that is destroyed while this solver is attached
    void forget_warm_start() {
        ws_ = nullptr;
        ws_mode_ = 0;
    }
    // views for the warm start
    const Dims &dev_dims() const { return d_; }
    const DevFields &dev_fields() const { return f_; }
    hipStream_t stream() const { return st_; }
    int beam_i() const { return istart_; }
    int beam_j() const { return jstart_; }
    std::map<std::string, KernelTimer> timers;

    // grouped runs (group.cpp): many mixed-mode solvers advance in lock step,
    // each kernel of an iteration one launch over all of them.
    // nonzero (with the error set) if this solver cannot take part
    int group_check() const;
    // every launch argument a recorded iteration holds by value
    std::string group_key() const;
    // the solve's start on the solver stream: warm start, the step's device
    // inputs, the loop reset; `ready` recorded after them
    int group_begin_solve(hipEvent_t ready);
    // one SIMPLE iteration's launches, recorded instead of issued
    int group_record_iteration(LaunchRecorder &rec);
    // the device's done flag of the step's loop
    int *group_done_flag() const { return &sc_->done; }
    // the solve's end once `done` (recorded after the grouped loop) has
    // passed: scalars, the step's status, the warm start's sample
    int group_end_solve(hipEvent_t done, float ms);

  private:
    double group_alenold_ = 0.0;
    amcfd_params p_{};
    std::vector<double> tool_, dtt_, heatt_;
    HostGrid hg_;
    Dims d_{};
    DevGeom g_{};
    DevFields f_{};
    Partials pr_{};
    DevScalars *sc_ = nullptr;       // device
    DevScalars *sc_host_ = nullptr;  // pinned mirror
    hipStream_t st_ = nullptr;
    // captured mixed loops: the second stream of the forked launches, its
    // fork / join events, the ghost values a launch run ahead keeps, whether
    // it ran, and the second set of the partials that alternate by iteration
    hipStream_t side_st_ = nullptr;
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    double *ghost_save_ = nullptr;
    int *ghost_spec_ = nullptr;
    double *energy_alt_ = nullptr, *bc_alt_ = nullptr;
    int mode_ = AMCFD_MODE_FP64;
    hipEvent_t step_ev_[2] = {nullptr, nullptr};   // brackets one solve_step
    std::vector<void *> allocs_;
    SweepPlan plan_dom_, plan_box_;
    unsigned long long *counter_ = nullptr;
    unsigned long long counter_base_ = 0;
    double *pass_buf_ = nullptr;   // four field-sized pass buffers of the line sweeps
    MxFields mx_{};                // mixed mode: FP32 coefficients, partials, second buffer
    // mixed mode, captured SIMPLE loop: a graph of graph_block_ iterations,
    // relaunched until the device reports the step done
    bool use_graph_ = true;
    int graph_block_ = 8;
    int box_blocks_ = 256;           // workgroups of the box kernels (fixed grids)
    static constexpr int max_box_blocks = 4096;   // flow partial slots reserved for them
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    struct GraphKey {
        int nk = -1, blocks = 0, box_blocks = 0;
        double delt = 0.0, scanvel = 0.0;
        std::string sched;
        int box_passes = 0, box_fused = -1, line_batch = 0, fork_asm = -1;
    } gkey_;
    StepDev *stepd_ = nullptr;       // device
    StepDev *steph_ = nullptr;       // pinned staging
    std::string mx_sched_ = "xxxx";   // mixed mode: order of the enthalpy line passes
    int mx_box_passes_ = 4;          // mixed mode: red-black x passes per box variable
    // mixed mode: the box passes of one variable in one persistent launch (1)
    // or one launch per half pass (0); the same results either way
    int box_fused_ = 1;
    int line_batch_ = 1;             // mixed mode: x-lines per workgroup of the enthalpy passes
    // captured mixed loops: the next iteration's boundary launch and far
    // assembly tiles run alongside the flow part (1) or after it (0); the
    // same results either way
    int fork_asm_ = 1;
    int fork_blocks_ = 0;           // workgroups of the far tiles' launch (0: one per tile)
    int fork_blocks_default_ = 0;   // its default, from the compute-unit count
    // the overlap a capture uses: none beside other solvers' streams
    // (amcfd_run_batch), where a stream waiting on another blocks the
    // hardware queue it shares with other solvers (measured 5x slower)
    int fork_key() const { return (fork_asm_ && !shared_device_) ? 1 + 2 * fork_blocks_ : 0; }
    double *box_bufs_ = nullptr;     // box_buf_passes_ pass buffers of ni*nj*nkf
    int box_buf_passes_ = 0;
    unsigned int *box_claim_ = nullptr;
    int ensure_box_buffers(int passes);
    int profile_ = 0;
    bool shared_device_ = false;
    std::vector<hipEvent_t> ev_pool_;
    std::vector<std::pair<std::string, std::pair<int, int>>> ev_pending_;
    size_t ev_next_ = 0;
    // host mirror of the reference's laser / driver state
    double timet_ = 0.0, scanvel_ = 0.0;
    double beamx_ = 0.0, beamy_ = 0.0, beamz_ = 0.0, oldx_ = 0.0, oldy_ = 0.0, oldz_ = 0.0;
    int powerindicator_ = 1, laser_init_ = 1, trackindx_ = 1, istart_ = 0, jstart_ = 0;
    int heat_it_ = 0, dt_state_ = 0, step_ = 0, niter_ = 0, itertot_ = 0;
    double alen_ = 0.0;
    bool steadylength_ = false;
    bool layer_overflow_ = false;            // a powder layer beyond the grid was requested
    std::vector<double> hg0_z_;              // z at set-up (the free surface's initial heights)
    double *fs_out_ = nullptr;               // free-surface search results (device, 6 values)
    double tracksect_ = 0.0;                 // time since the last extracted section
    bool section_sci_ = false;               // the section stream has switched to scientific
    double *xold_ = nullptr, *yold_ = nullptr;   // node coordinates before a mesh move
    Box box_{};
    WarmStart *ws_ = nullptr;
    int ws_mode_ = 0;
    int last_solve_rc_ = AMCFD_I_MAXIT;   // the previous solve's status (none yet: not converged)
    int ws_applied_ = 0, ws_guarded_ = 0;  // steps started from the guess / kept on the plain start
    int warm_start_begin();
    int warm_start_end();

    void set_active_nk(int nk);
    int alloc(double **ptr, size_t n);
    int upload(double *dst, const std::vector<double> &src);
    void laser_position();
    int launch_laser_map();
    int move_mesh();
  public:
    // energy-based free surface at an output time (monitor appended if path)
    int free_surface(const char *monitor_path);
    // PrintManager::outputSection: accumulate dt, write when due or at the end
    int output_section(const char *base, int finish, int *written);
    struct FsMonitor {
        double xmid, fmid, fimin, fimax, vd;
        int niter, warn;
    } fs_monitor_{};
  private:
    int upload_geometry();
    SourceParams source_params();
    int sweep(double *phi, SweepPlan &plan, int ibeg, int iend, int jbeg, int jend, int kbeg, int kend, int guard,
              const char *timer);
    int ensure_plan(SweepPlan &plan, int kbeg, int kend, int jbeg, int jend);
    int read_scalars();
    int flow_iteration();
    int solve_step_mixed();
    int solve_step_mixed_graph();
    int solve_step_fp64();
    // one mixed-mode SIMPLE iteration; ctl: every decision on the device
    // (flags in DevScalars, the step's inputs in stepd_), for capture.
    // it: the iteration's place in a captured block (its partial buffers
    // alternate); ahead: its boundary launch and the tiles of its assembly
    // away from the melt-pool box were enqueued by the iteration before, on
    // side_st_; fork_next: enqueue those of the next iteration on side_st_,
    // alongside this iteration's flow part
    int enqueue_mixed_iteration(bool ctl, int it = 0, bool ahead = false, bool fork_next = false);
    int capture_mixed_graph();
    void release_graph();
    int energy_sweeps_mixed(const int *skip);
    int allocf(float **ptr, size_t n);
    void tic(const char *name, int *slot);
    void toc(const char *name, int slot);
    void collect_timers();
    double *field_ptr(const char *name);
};

// grouped lock-step runs of mixed-mode solvers (group.cpp)
int run_grouped(const std::vector<Solver *> &s, int max_steps, int *steps_done);

// output writers (output.cpp)
int write_vts(Solver &s, const char *path, const char *const *names, int nnames);
int write_pvd(const char *path, int n, const double *times, const char *const *files);
int append_residuals(Solver &s, const char *path);
int append_final_geometry(Solver &s, const char *path);
int append_free_surface_monitor(Solver &s, const char *path);
int write_section(Solver &s, const char *base, bool &scientific_state);

}  // namespace amcfd
