// Host driver of the device-resident SIMPLE solve.
//
// The time-step and iteration structure is the reference's:
//   step:      dt table, timet += dt, updateLaser       src/main.cpp:125-141
//   solve:     CFDSolverManager::iterativeLoopSolve     CFDSolverManager.cpp:595-722
//   push-back: CFDSolverManager::updateSolutions        CFDSolverManager.cpp:2625-2687
// All field arithmetic runs in HIP kernels on one stream; the host keeps the
// scalar control state (laser position, heat-source table row, convergence
// test) and reads back one small scalar block per SIMPLE iteration.

#include "host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace amcfd {

#define HIPCALL(x)                                                                     \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) return set_error(AMCFD_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

void Solver::release_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    gexec_ = nullptr;
    graph_ = nullptr;
    gkey_ = GraphKey();
}

Solver::~Solver() {
    // teardown: errors here have no caller to go to
    if (st_) (void)hipStreamSynchronize(st_);
    if (ws_) ws_detach(ws_, this);
    release_graph();
    if (steph_) (void)hipHostFree(steph_);
    for (void *p : allocs_) (void)hipFree(p);
    if (sc_host_) (void)hipHostFree(sc_host_);
    for (auto &e : ev_pool_) (void)hipEventDestroy(e);
    for (auto &e : step_ev_)
        if (e) (void)hipEventDestroy(e);
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (side_st_) {
        (void)hipStreamSynchronize(side_st_);
        (void)hipStreamDestroy(side_st_);
    }
    if (st_) (void)hipStreamDestroy(st_);
}

int Solver::alloc(double **ptr, size_t n) {
    void *p = nullptr;
    HIPCALL(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)));
    HIPCALL(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(double), st_));
    allocs_.push_back(p);
    *ptr = (double *)p;
    return AMCFD_OK;
}

int Solver::allocf(float **ptr, size_t n) {
    void *p = nullptr;
    HIPCALL(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(float)));
    HIPCALL(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(float), st_));
    allocs_.push_back(p);
    *ptr = (float *)p;
    return AMCFD_OK;
}

int Solver::set_graph(int on, int block, int box_blocks) {
    if (block < 1 || block > 64) return set_error(AMCFD_E_ARG, "graph block must be 1..64 iterations");
    if (box_blocks < 1 || box_blocks > max_box_blocks)
        return set_error(AMCFD_E_ARG, "box workgroups must be 1.." + std::to_string(max_box_blocks));
    use_graph_ = on != 0;
    graph_block_ = block;
    box_blocks_ = box_blocks;
    release_graph();
    return AMCFD_OK;
}

int Solver::ensure_box_buffers(int passes) {
    if (box_buf_passes_ >= passes) return AMCFD_OK;
    // outside any capture: the buffers are allocated before a graph is taken
    const size_t nn = (size_t)d_.ni * d_.nj * d_.nkf;
    double *p = nullptr;
    int st = alloc(&p, nn * passes);
    if (st) return st;
    box_bufs_ = p;   // an earlier, smaller set stays allocated until the solver goes
    box_buf_passes_ = passes;
    return AMCFD_OK;
}

int Solver::set_box_fused(int on) {
    box_fused_ = on != 0;
    release_graph();
    return box_fused_ ? ensure_box_buffers(mx_box_passes_) : AMCFD_OK;
}

int Solver::set_assembly_overlap(int on) {
    if (on < 0) return set_error(AMCFD_E_ARG, "overlap: 0, 1 or a workgroup count");
    fork_asm_ = on != 0;
    fork_blocks_ = on > 1 ? on : fork_blocks_default_;
    return AMCFD_OK;
}

int Solver::set_line_batch(int lines) {
    if (lines < 1 || lines > 64) return set_error(AMCFD_E_ARG, "line batch must be 1..64");
    line_batch_ = lines;
    return AMCFD_OK;
}

int Solver::set_mixed_schedule(const char *sched, int box_passes) {
    if (!sched || !*sched) return set_error(AMCFD_E_ARG, "empty pass schedule");
    int yz = 0;
    for (const char *c = sched; *c; ++c) {
        if (*c != 'x' && *c != 'y' && *c != 'z') return set_error(AMCFD_E_ARG, "passes are 'x', 'y' or 'z'");
        yz += *c != 'x';
    }
    // the y/z passes alternate between the field and its second buffer; an
    // even count ends in the field itself
    if (yz % 2) return set_error(AMCFD_E_ARG, "the schedule needs an even number of y/z passes");
    if (box_passes < 1 || box_passes > 16) return set_error(AMCFD_E_ARG, "box passes must be 1..16");
    if (yz) {   // the y/z passes keep a whole line's forward values in LDS
        const int L = std::max(d_.njm2, d_.nkf - 2);
        if (L > mixed_max_yz())
            return set_error(AMCFD_E_CAPACITY, "y/z passes take lines of at most " + std::to_string(mixed_max_yz()) +
                                                   " cells");
        int dev = 0, lds_max = 0;
        HIPCALL(hipGetDevice(&dev));
        HIPCALL(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (mixed_yz_lds(L) > (size_t)lds_max)
            return set_error(AMCFD_E_CAPACITY, "y- or z-lines too long for the device's LDS");
    }
    mx_sched_ = sched;
    mx_box_passes_ = box_passes;
    release_graph();
    return box_fused_ ? ensure_box_buffers(mx_box_passes_) : AMCFD_OK;
}

int Solver::upload(double *dst, const std::vector<double> &src) {
    HIPCALL(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice, st_));
    return AMCFD_OK;
}

void Solver::set_active_nk(int nk) {
    d_.nk = nk;
    d_.nkm1 = nk - 1;
    d_.nkm2 = nk - 2;
}

int Solver::create(const amcfd_params &p, int mode) {
    if (mode != AMCFD_MODE_FP64 && mode != AMCFD_MODE_SCAN && mode != AMCFD_MODE_MIXED)
        return set_error(AMCFD_E_ARG, "unknown solver mode");
    const bool mixed = mode == AMCFD_MODE_MIXED;
    mode_ = mode;
    p_ = p;
    if (p.n_toolpath > 0) tool_.assign(p.toolpath, p.toolpath + 6 * p.n_toolpath);
    if (p.n_dt_table > 0) dtt_.assign(p.dt_table, p.dt_table + 3 * p.n_dt_table);
    if (p.n_heat_table > 0) heatt_.assign(p.heat_table, p.heat_table + 6 * p.n_heat_table);
    p_.toolpath = tool_.empty() ? nullptr : tool_.data();
    p_.dt_table = dtt_.empty() ? nullptr : dtt_.data();
    p_.heat_table = heatt_.empty() ? nullptr : heatt_.data();
    if (p.ni < 4 || p.nj < 4 || p.nk < 4) return set_error(AMCFD_E_ARG, "grid too small");
    if (!mixed && p.ni - 2 > sweep_max_line())
        return set_error(AMCFD_E_CAPACITY, "x-lines longer than " + std::to_string(sweep_max_line()) + " cells");
    if (mixed && p.ni - 2 > mixed_max_line())
        return set_error(AMCFD_E_CAPACITY, "x-lines longer than " + std::to_string(mixed_max_line()) +
                                               " cells in the mixed mode");
    if (mixed && p.species)
        return set_error(AMCFD_E_UNSUPPORTED, "species transport runs in the FP64 and scan modes only");
    if (p.nj >= 16384 || p.nk >= 16384) return set_error(AMCFD_E_CAPACITY, "more than 16383 nodes in y or z");
    if (p.maxit < 1) return set_error(AMCFD_E_ARG, "maxit must be >= 1");
    if (p.acpa == 0.0) return set_error(AMCFD_E_ARG, "acpa must be nonzero (solid H->T divides by it)");

    int dev = 0;
    HIPCALL(hipGetDevice(&dev));
    if (mixed) {   // the captured loop's launch shapes for this size, from the tuning table if there is one
        tuning_lookup((long long)(p.ni - 2) * (p.nj - 2) * (p.nk - 2), &graph_block_, &box_blocks_, &box_fused_,
                      &line_batch_, &fork_asm_);
        box_blocks_ = std::min(box_blocks_, max_box_blocks);
        // the far assembly tiles beside the flow part: about one and a half
        // workgroups per compute unit, which leaves room on every unit for
        // the flow part's launches (measured best on the 400x160x80 track)
        int ncu = 0;
        HIPCALL(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        fork_blocks_ = std::max(8, 3 * ncu / 2);
        fork_blocks_default_ = fork_blocks_;
    }
    HIPCALL(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    HIPCALL(hipEventCreate(&step_ev_[0]));
    HIPCALL(hipEventCreate(&step_ev_[1]));

    d_.ni = p.ni;
    d_.nj = p.nj;
    d_.nkf = p.nk;
    d_.nim1 = p.ni - 1;
    d_.nim2 = p.ni - 2;
    d_.njm1 = p.nj - 1;
    d_.njm2 = p.nj - 2;
    d_.sj = (size_t)p.ni;
    d_.sk = (size_t)p.ni * p.nj;
    set_active_nk(p.nk);

    hg_.build(p_);
    hg0_z_ = hg_.z;
    const size_t nn = (size_t)d_.ni * d_.nj * d_.nkf;
    const int n1 = d_.ni + d_.nj + d_.nkf + 8;
    int st;
    // geometry
    double **g1[] = {&g_.x, &g_.y, &g_.z, &g_.xu, &g_.yv, &g_.zw, &g_.dxpwinv, &g_.dypsinv, &g_.dzpbinv,
                     &g_.fracx, &g_.fracy, &g_.fracz};
    const std::vector<double> *h1[] = {&hg_.x, &hg_.y, &hg_.z, &hg_.xu, &hg_.yv, &hg_.zw, &hg_.dxpwinv,
                                       &hg_.dypsinv, &hg_.dzpbinv, &hg_.fracx, &hg_.fracy, &hg_.fracz};
    for (int a = 0; a < 12; ++a) {
        if ((st = alloc(g1[a], n1))) return st;
        if ((st = upload(*g1[a], *h1[a]))) return st;
    }
    if (p_.movingmesh) {
        if ((st = alloc(&xold_, n1))) return st;
        if ((st = alloc(&yold_, n1))) return st;
    }
    double **g2[] = {&g_.areaij, &g_.areauij, &g_.areavij, &g_.areaik, &g_.areawik, &g_.areauik,
                     &g_.areajk, &g_.areavjk, &g_.areawjk};
    const std::vector<double> *h2[] = {&hg_.areaij, &hg_.areauij, &hg_.areavij, &hg_.areaik, &hg_.areawik,
                                       &hg_.areauik, &hg_.areajk, &hg_.areavjk, &hg_.areawjk};
    for (int a = 0; a < 9; ++a) {
        if ((st = alloc(g2[a], h2[a]->size()))) return st;
        if ((st = upload(*g2[a], *h2[a]))) return st;
    }
    // fields
    double **fl[] = {&f_.enthalpy, &f_.temp, &f_.fracl, &f_.uvel, &f_.vvel, &f_.wvel, &f_.pressure, &f_.pp,
                     &f_.rho, &f_.diff, &f_.vis, &f_.hnot, &f_.fraclnot, &f_.unot, &f_.vnot, &f_.wnot,
                     &f_.sourceinput, &f_.dux, &f_.dvy, &f_.dwz, &f_.csfrac, &f_.solfrac, &f_.mpphase,
                     &f_.mpliqphase, &f_.ap, &f_.an, &f_.as, &f_.ae, &f_.aw, &f_.at, &f_.ab, &f_.su, &f_.sp,
                     &f_.apnot, &f_.lpr, &f_.lden, &f_.ly};
    for (auto *q : fl)
        if ((st = alloc(q, nn))) return st;
    if ((st = alloc(&f_.heatin, (size_t)d_.ni * d_.nj))) return st;
    if (p_.species) {   // CFDSolverManager.cpp:474-479
        double **sp3[] = {&f_.conc, &f_.concnot, &f_.concl};
        for (auto *q : sp3) {
            if ((st = alloc(q, nn))) return st;
            if ((st = upload(*q, std::vector<double>(nn, p_.initconcentration)))) return st;
        }
    }
    if (p_.energyfreesurface) {   // FreeSurfaceManager::initializeFreeSurface, :35-67
        if ((st = alloc(&f_.fi, (size_t)d_.ni * d_.nj))) return st;
        if ((st = alloc(&f_.zzm, (size_t)d_.ni * d_.nj))) return st;
        if ((st = alloc(&f_.zr, nn))) return st;
        if ((st = alloc(&f_.tempr, nn))) return st;
        std::vector<double> zr(nn);
        for (size_t c = 0; c < nn; ++c) zr[c] = hg_.z[c / ((size_t)d_.ni * d_.nj)];
        // tempr starts as the initial temperature, uniform preheat (CFDSolverManager.cpp:429-447)
        if ((st = upload(f_.zr, zr)) || (st = upload(f_.tempr, std::vector<double>(nn, p_.temppreheat)))) return st;
        if ((st = alloc(&fs_out_, 8))) return st;
    }
    if (p_.solidification) {   // zero until the first molten step (CFDSolverManager.cpp:469-473)
        if ((st = alloc(&f_.grad, nn))) return st;
        if ((st = alloc(&f_.rate, nn))) return st;
    }
    InitialState is;
    initial_fields(p_, hg_, is);
    if (mx_.halt && (st = upload(mx_.halt, is.enthalpy))) return st;   // finite ghost values
    if ((st = upload(f_.enthalpy, is.enthalpy)) || (st = upload(f_.hnot, is.hnot)) ||
        (st = upload(f_.temp, is.temp)) || (st = upload(f_.rho, is.rho)) || (st = upload(f_.vis, is.vis)) ||
        (st = upload(f_.diff, is.diff)) || (st = upload(f_.csfrac, is.csfrac)))
        return st;
    // partial-sum buffers sized for the whole interior
    energy_partial_sizes(d_, &pr_.nblk_e, &pr_.nrow_e, &pr_.nblk_t);
    Box whole{1, d_.nim2, 1, d_.njm2, 2};
    // the flow partials hold one slot per workgroup of the box kernels: the
    // box's own tiles (at most the whole interior's), or the fixed
    // workgroup count of the mixed mode's launches
    int nf = std::max(flow_partial_blocks(whole, d_) + 64, mixed ? max_box_blocks : 0);
    if ((st = alloc(&pr_.energy, (size_t)2 * pr_.nblk_e))) return st;
    if ((st = alloc(&pr_.species, (size_t)2 * pr_.nblk_e))) return st;
    // block correction: 4 column sums, each ni long
    if ((st = alloc(&pr_.blk, (size_t)4 * d_.ni + 8))) return st;
    if ((st = alloc(&pr_.h2t, (size_t)10 * pr_.nblk_t))) return st;
    if ((st = alloc(&pr_.flow, (size_t)4 * nf * 3))) return st;
    pr_.nblk_b = (d_.njm2 * d_.nim2 + d_.nkm2 * d_.njm2 + d_.nkm2 * d_.nim2 + 255) / 256;
    if ((st = alloc(&pr_.bc, (size_t)pr_.nblk_b))) return st;
    if ((st = alloc(&pr_.delphi, (size_t)d_.ni + 2))) return st;
    if ((st = alloc(&pr_.laser, (size_t)d_.ni * d_.nj))) return st;
    void *v = nullptr;
    HIPCALL(hipMalloc(&v, sizeof(DevScalars)));
    allocs_.push_back(v);
    sc_ = (DevScalars *)v;
    HIPCALL(hipMemsetAsync(sc_, 0, sizeof(DevScalars), st_));
    HIPCALL(hipHostMalloc((void **)&sc_host_, sizeof(DevScalars), hipHostMallocDefault));
    std::memset(sc_host_, 0, sizeof(DevScalars));
    HIPCALL(hipMalloc(&v, sizeof(unsigned long long)));
    allocs_.push_back(v);
    counter_ = (unsigned long long *)v;
    HIPCALL(hipMemsetAsync(counter_, 0, sizeof(unsigned long long), st_));
    if (!mixed) {   // four pass buffers for the Gauss-Seidel line sweeps
        if ((st = alloc(&pass_buf_, 4 * nn))) return st;
    } else {
        float **c32[] = {&mx_.ap, &mx_.an, &mx_.as, &mx_.ae, &mx_.aw, &mx_.at, &mx_.ab};
        for (auto *q : c32)
            if ((st = allocf(q, nn))) return st;
        // the energy equation's source of its own (the momentum equations'
        // use f.su while the next iteration's assembly runs alongside them)
        if ((st = alloc(&mx_.su, nn))) return st;
        if ((st = alloc(&energy_alt_, (size_t)2 * pr_.nblk_e))) return st;
        if ((st = alloc(&bc_alt_, (size_t)pr_.nblk_b))) return st;
        {
            Dims full = d_;   // the ghost layer of the full z extent
            full.nk = d_.nkf;
            full.nkm1 = d_.nkf - 1;
            full.nkm2 = d_.nkf - 2;
            if ((st = alloc(&ghost_save_, bc_ghost_values(full)))) return st;
            void *v = nullptr;
            HIPCALL(hipMalloc(&v, sizeof(int)));
            allocs_.push_back(v);
            ghost_spec_ = (int *)v;
            HIPCALL(hipMemsetAsync(ghost_spec_, 0, sizeof(int), st_));
        }
        HIPCALL(hipStreamCreateWithFlags(&side_st_, hipStreamNonBlocking));
        HIPCALL(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
        HIPCALL(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
        Dims full = d_;   // partials sized for the full z extent
        full.nk = d_.nkf;
        full.nkm1 = d_.nkf - 1;
        full.nkm2 = d_.nkf - 2;
        if ((st = alloc(&mx_.blkpart, (size_t)4 * mixed_tile_count(full) * d_.ni))) return st;
        if ((st = alloc(&mx_.blkfold, (size_t)4 * mixed_blk_split() * d_.ni))) return st;
        if ((st = alloc(&mx_.delphi, (size_t)d_.ni + 2))) return st;
        if ((st = alloc(&mx_.halt, nn))) return st;
        {
            void *v = nullptr;
            HIPCALL(hipMalloc(&v, sizeof(unsigned int)));
            allocs_.push_back(v);
            box_claim_ = (unsigned int *)v;
            HIPCALL(hipMemsetAsync(box_claim_, 0, sizeof(unsigned int), st_));
        }
        if (box_fused_ && (st = ensure_box_buffers(mx_box_passes_))) return st;
        void *v = nullptr;
        HIPCALL(hipMalloc(&v, sizeof(StepDev)));
        allocs_.push_back(v);
        stepd_ = (StepDev *)v;
        HIPCALL(hipHostMalloc((void **)&steph_, sizeof(StepDev), hipHostMallocDefault));
        std::memset(steph_, 0, sizeof(StepDev));
    }

    scanvel_ = p_.scanvel;
    set_active_nk(p_.nk_active0);
    if (p_.steady) {   // main.cpp:105-106
        beamx_ = p_.xstart;
        int s2 = launch_laser_map();
        if (s2) return s2;
    }
    HIPCALL(hipStreamSynchronize(st_));
    return AMCFD_OK;
}

// ---------------------------------------------------------------------------
// laser: BoundCondManager.cpp:68-275
// ---------------------------------------------------------------------------
void Solver::laser_position() {
    if (!p_.udtoolpath) {
        oldx_ = beamx_;
        if (timet_ <= p_.lasertime) {
            beamx_ = p_.xstart + timet_ * scanvel_;
            if (laser_init_) {
                oldx_ = beamx_;
                laser_init_ = 0;
            }
        } else {
            beamx_ = p_.xstart + p_.lasertime * scanvel_;
            powerindicator_ = 0;
        }
        beamy_ = p_.ystart;
        return;
    }
    const double small = 1.0e-8;
    double tx = beamx_, ty = beamy_, tz = beamz_, dty = 0.0;
    oldx_ = beamx_;
    oldy_ = beamy_;
    oldz_ = beamz_;
    int oldpi = powerindicator_;
    for (int r = 1; r < p_.n_toolpath; ++r) {
        const double *b = &tool_[6 * r], *a = &tool_[6 * (r - 1)];
        if (timet_ <= b[0] + small) {
            double ratio = (timet_ - a[0]) / (b[0] - a[0]);
            double dtx = b[1] - a[1];
            dty = b[2] - a[2];
            double dtz = b[3] - a[3];
            tx = ratio * dtx + a[1];
            ty = ratio * dty + a[2];
            tz = ratio * dtz + a[3];
            powerindicator_ = (int)b[4];
            if (laser_init_) {
                oldx_ = a[1];
                oldy_ = a[2];
                oldz_ = a[3];
                oldpi = (int)a[4];
                laser_init_ = 0;
            }
            break;
        }
    }
    beamx_ = tx;
    beamy_ = ty;
    beamz_ = tz;
    double dx = beamx_ - oldx_, dy = beamy_ - oldy_, dz = beamz_ - oldz_;
    scanvel_ = std::sqrt(dx * dx + dy * dy + dz * dz) / p_.delt;
    if (oldpi - powerindicator_ > 0 && dty > small) trackindx_++;
    if (powerindicator_ - oldpi > 0 && p_.powderbed) {
        trackindx_++;
        if (d_.nk + p_.ncvzlayer <= d_.nkf) set_active_nk(d_.nk + p_.ncvzlayer);
        else layer_overflow_ = true;
    }
}

static int nearest(const std::vector<double> &c, int nlast, double pos) {
    int out = 0;
    for (int i = 1; i < nlast; ++i) {
        if (pos <= c[i]) break;
        out = i;
    }
    if (std::fabs(pos - c[out + 1]) < std::fabs(pos - c[out])) out++;
    return out;
}

// Gaussian surface flux map, BoundCondManager.cpp:128-159, evaluated on the
// device (k_laser_map); the nearest grid indices stay on the host.
int Solver::launch_laser_map() {
    istart_ = nearest(hg_.x, d_.nim1, beamx_);
    jstart_ = nearest(hg_.y, d_.njm1, beamy_);
    LaserParams q;
    q.beamx = beamx_;
    q.beamy = beamy_;
    q.alasfact = p_.alasfact;
    q.rb2 = p_.alasrb * p_.alasrb;
    q.varlas = p_.alaspow * p_.alaseta;
    q.peakhin = p_.alasfact * q.varlas / (M_PI * q.rb2);
    q.pind = (double)powerindicator_;
    amcfd::launch_laser_map(st_, d_, g_, f_, pr_, q, sc_);
    HIPCALL(hipGetLastError());
    return AMCFD_OK;
}

int Solver::begin_step() {
    if (!dtt_.empty()) {
        if (timet_ > dtt_[3 * dt_state_] && dt_state_ < p_.n_dt_table - 1) dt_state_ += 1;
        p_.delt = dtt_[3 * dt_state_ + 1];
        p_.outtime = dtt_[3 * dt_state_ + 2];
    }
    timet_ += p_.delt;
    ++step_;
    if (!p_.steady) {
        laser_position();
        if (layer_overflow_)
            return set_error(AMCFD_E_CAPACITY, "the toolpath starts more powder layers than *POWDER_BED_MODE declares");
        int st = launch_laser_map();
        if (st) return st;
    }
    if (p_.movingmesh) return move_mesh();   // main.cpp:144-145
    return AMCFD_OK;
}

// FreeSurfaceManager::execute (FreeSurfaceManager.cpp:72-220) at an output
// time: the added droplet volume on the host, the multiplier search and the
// profile on the device; appends the monitor block when path is given
int Solver::free_surface(const char *monitor_path) {
    if (!p_.energyfreesurface) return set_error(AMCFD_E_ARG, "the deck has no *TRACK_ENERGY_BASED_FREE_SURFACE");
    const double vdp = 2e-9, startdeltn = 0.0;
    double melttime = p_.delt * (1.0 + startdeltn), vd;
    if (sc_host_->tpeak <= p_.tsolid) {
        vd = 0.0;
        melttime = 0.0;
    } else {
        vd = vdp * (timet_ - melttime - startdeltn * p_.delt);
        if (powerindicator_ < 1) vd = vdp * (p_.lasertime - melttime);
    }
    launch_free_surface(st_, d_, g_, f_, p_, vd, fs_out_);
    HIPCALL(hipGetLastError());
    double r[6];
    HIPCALL(hipMemcpyAsync(r, fs_out_, sizeof r, hipMemcpyDeviceToHost, st_));
    HIPCALL(hipStreamSynchronize(st_));
    fs_monitor_ = {r[0], r[1], r[2], r[3], vd, (int)r[4], (int)r[5]};
    if (monitor_path) return append_free_surface_monitor(*this, monitor_path);
    return AMCFD_OK;
}

// PrintManager::outputSection (PrintManager.cpp:538-624): called every step
// after the solve, and once more at the end with finish set
int Solver::output_section(const char *base, int finish, int *written) {
    *written = 0;
    if (!p_.outputsection) return set_error(AMCFD_E_ARG, "the deck has no *OUTPUT_SECTION");
    tracksect_ += p_.delt;
    if (tracksect_ < p_.toutsect && !finish) return AMCFD_OK;
    int st = write_section(*this, base, section_sci_);
    if (st) return st;
    tracksect_ = 0.0;
    *written = 1;
    return AMCFD_OK;
}

// all grid arrays, host -> device (the z extent may be the active one only,
// the rest keeps its earlier values on both sides)
int Solver::upload_geometry() {
    double *g1[] = {g_.x, g_.y, g_.z, g_.xu, g_.yv, g_.zw, g_.dxpwinv, g_.dypsinv, g_.dzpbinv,
                    g_.fracx, g_.fracy, g_.fracz, g_.areaij, g_.areauij, g_.areavij, g_.areaik,
                    g_.areawik, g_.areauik, g_.areajk, g_.areavjk, g_.areawjk};
    const std::vector<double> *h1[] = {&hg_.x, &hg_.y, &hg_.z, &hg_.xu, &hg_.yv, &hg_.zw, &hg_.dxpwinv,
                                       &hg_.dypsinv, &hg_.dzpbinv, &hg_.fracx, &hg_.fracy, &hg_.fracz,
                                       &hg_.areaij, &hg_.areauij, &hg_.areavij, &hg_.areaik, &hg_.areawik,
                                       &hg_.areauik, &hg_.areajk, &hg_.areavjk, &hg_.areawjk};
    int st;
    for (int a = 0; a < 21; ++a)
        if ((st = upload(g1[a], *h1[a]))) return st;
    return AMCFD_OK;
}

// CFDSolverManager::generateGrid after set-up (CFDSolverManager.cpp:60-124,
// 345-346): the zones either side of the moving one absorb the beam's
// displacement since the last step, the grid is rebuilt over the active
// extent, and the solution is remapped onto it on the device.  Until the
// pool length is steady the reference picks the y zones with the x index.
int Solver::move_mesh() {
    const std::vector<double> px = hg_.x, py = hg_.y;
    const double dx = beamx_ - oldx_, dy = beamy_ - oldy_;
    const int dirx = (dx < 0) ? -1 : 1, diry = (dy < 0) ? -1 : 1;
    const int ix = p_.xmoveindex, iy = p_.ymoveindex;
    p_.xzone[ix - 1] += dx;
    p_.xzone[ix + 1] -= dx;
    const int iyz = steadylength_ ? iy : ix;
    p_.yzone[iyz - 1] += dy;
    p_.yzone[iyz + 1] -= dy;
    hg_.regenerate(p_, d_.nk);
    // the laser map queued above still reads the old grid
    HIPCALL(hipStreamSynchronize(st_));
    int st;
    if ((st = upload_geometry()) || (st = upload(xold_, px)) || (st = upload(yold_, py))) return st;
    launch_remap(st_, d_, f_, g_, xold_, yold_, dirx, diry, p_.species);
    HIPCALL(hipGetLastError());
    // the host vectors are rewritten at the next move
    HIPCALL(hipStreamSynchronize(st_));
    return AMCFD_OK;
}

// volumetric-source parameters, CFDSolverManager.cpp:1686-1718 (the table
// row advances at most once per call, i.e. once per SIMPLE iteration)
SourceParams Solver::source_params() {
    SourceParams s{0, 0, 0, 0, 0, beamx_, beamy_, powerindicator_};
    if (p_.volheatsource && !p_.volheatsourcetable) {
        s.avolpow = p_.avolpow;
        s.apowseta = p_.apowseta;
        s.heatthick = p_.heatthick;
        s.heatrb = p_.heatrb;
        s.avolfact = p_.avolfact;
    } else if (!heatt_.empty()) {
        if (timet_ > heatt_[6 * heat_it_] && heat_it_ < p_.n_heat_table - 1) heat_it_ += 1;
        const double *r = &heatt_[6 * heat_it_];
        s.avolpow = r[1];
        s.apowseta = r[2];
        s.heatthick = r[3];
        s.heatrb = r[4];
        s.avolfact = r[5];
    }
    return s;
}

// ---------------------------------------------------------------------------
// profiling helpers (events on the solver stream)
// ---------------------------------------------------------------------------
void Solver::tic(const char *name, int *slot) {
    *slot = -1;
    if (!profile_) return;
    if (profile_ == 2 && std::strcmp(name, "sweep_h") != 0) return;   // the enthalpy sweep alone
    if (ev_next_ + 2 > ev_pool_.size()) {
        for (int a = 0; a < 64; ++a) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            ev_pool_.push_back(e);
        }
    }
    if (hipEventRecord(ev_pool_[ev_next_], st_) != hipSuccess) return;
    *slot = (int)ev_next_;
    ev_next_ += 2;
}

void Solver::toc(const char *name, int slot) {
    if (slot < 0) return;
    if (hipEventRecord(ev_pool_[slot + 1], st_) != hipSuccess) return;
    ev_pending_.push_back({name, {slot, slot + 1}});
}

void Solver::collect_timers() {
    for (auto &pe : ev_pending_) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev_pool_[pe.second.first], ev_pool_[pe.second.second]) == hipSuccess) {
            KernelTimer &t = timers[pe.first];
            t.launches += 1;
            t.total_ms += ms;
        }
    }
    ev_pending_.clear();
    ev_next_ = 0;
}

// ---------------------------------------------------------------------------
// line sweeps
// ---------------------------------------------------------------------------
int Solver::ensure_plan(SweepPlan &plan, int kbeg, int kend, int jbeg, int jend) {
    if (plan.kbeg == kbeg && plan.kend == kend && plan.jbeg == jbeg && plan.jend == jend && plan.tasks) return 0;
    std::vector<int> tasks;
    int width = 0;
    build_sweep_tasks(kbeg, kend, jbeg, jend, tasks, width);
    if ((int)tasks.size() > plan.cap) {
        if (plan.tasks) {
            HIPCALL(hipStreamSynchronize(st_));
            HIPCALL(hipFree(plan.tasks));
        }
        void *v = nullptr;
        HIPCALL(hipMalloc(&v, sizeof(int) * std::max<size_t>(tasks.size(), 1)));
        plan.tasks = (int *)v;
        plan.cap = (int)tasks.size();
    }
    if (!tasks.empty()) {
        // the copy is ordered on the stream after every earlier sweep
        HIPCALL(hipMemcpyAsync(plan.tasks, tasks.data(), tasks.size() * sizeof(int), hipMemcpyHostToDevice, st_));
        HIPCALL(hipStreamSynchronize(st_));
    }
    plan.ntasks = (int)tasks.size();
    plan.width = width;
    plan.kbeg = kbeg;
    plan.kend = kend;
    plan.jbeg = jbeg;
    plan.jend = jend;
    return 0;
}

int Solver::sweep(double *phi, SweepPlan &plan, int ibeg, int iend, int jbeg, int jend, int kbeg, int kend,
                  int guard, const char *timer) {
    int st = ensure_plan(plan, kbeg, kend, jbeg, jend);
    if (st) return st;
    if (plan.ntasks == 0 || iend < ibeg) return 0;
    // alone, twice the wavefront width of workgroups keeps the next tasks'
    // coefficients in flight; beside other solvers the device's workgroup
    // slots are shared, and one per wavefront task is enough
    int nblocks = std::min(plan.ntasks, std::max((shared_device_ ? 1 : 2) * plan.width, 1));
    nblocks = std::min(nblocks, 2048);
    launch_line_prep(st_, d_, f_, pass_buf_, (size_t)d_.ni * d_.nj * d_.nkf, plan, ibeg, iend, guard);
    int slot;
    tic(timer, &slot);
    launch_line_sweeps(st_, d_, f_, phi, pass_buf_, (size_t)d_.ni * d_.nj * d_.nkf, plan, ibeg, iend, guard,
                       mode_ == AMCFD_MODE_SCAN, counter_, counter_base_, sc_, nblocks, shared_device_);
    HIPCALL(hipGetLastError());
    toc(timer, slot);
    counter_base_ += (unsigned long long)plan.ntasks + (unsigned long long)nblocks;
    return 0;
}

int Solver::copy_state_from(const Solver &o) {
    if (o.d_.ni != d_.ni || o.d_.nj != d_.nj || o.d_.nkf != d_.nkf || o.p_.species != p_.species ||
        o.p_.energyfreesurface != p_.energyfreesurface || o.p_.solidification != p_.solidification)
        return set_error(AMCFD_E_ARG, "the solvers do not run the same grid and keywords");
    const size_t nn = (size_t)d_.ni * d_.nj * d_.nkf;
    // the state a step starts from; coefficient arrays are rebuilt by every
    // iteration
    double *const dst[] = {f_.enthalpy, f_.temp, f_.fracl, f_.uvel, f_.vvel, f_.wvel, f_.pressure, f_.pp,
                           f_.rho, f_.diff, f_.vis, f_.hnot, f_.fraclnot, f_.unot, f_.vnot, f_.wnot,
                           f_.sourceinput, f_.dux, f_.dvy, f_.dwz, f_.csfrac, f_.solfrac, f_.mpphase,
                           f_.mpliqphase, f_.conc, f_.concnot, f_.concl, f_.grad, f_.rate, f_.zr, f_.tempr};
    const double *const src[] = {o.f_.enthalpy, o.f_.temp, o.f_.fracl, o.f_.uvel, o.f_.vvel, o.f_.wvel,
                                 o.f_.pressure, o.f_.pp, o.f_.rho, o.f_.diff, o.f_.vis, o.f_.hnot, o.f_.fraclnot,
                                 o.f_.unot, o.f_.vnot, o.f_.wnot, o.f_.sourceinput, o.f_.dux, o.f_.dvy, o.f_.dwz,
                                 o.f_.csfrac, o.f_.solfrac, o.f_.mpphase, o.f_.mpliqphase, o.f_.conc,
                                 o.f_.concnot, o.f_.concl, o.f_.grad, o.f_.rate, o.f_.zr, o.f_.tempr};
    HIPCALL(hipStreamSynchronize(o.st_));
    for (size_t a = 0; a < sizeof(dst) / sizeof(dst[0]); ++a)
        if (dst[a] && src[a]) HIPCALL(hipMemcpyAsync(dst[a], src[a], nn * sizeof(double), hipMemcpyDeviceToDevice, st_));
    if (mx_.halt) HIPCALL(hipMemcpyAsync(mx_.halt, o.f_.enthalpy, nn * sizeof(double), hipMemcpyDeviceToDevice, st_));
    HIPCALL(hipMemcpyAsync(f_.heatin, o.f_.heatin, (size_t)d_.ni * d_.nj * sizeof(double), hipMemcpyDeviceToDevice, st_));
    if (f_.fi) {
        HIPCALL(hipMemcpyAsync(f_.fi, o.f_.fi, (size_t)d_.ni * d_.nj * sizeof(double), hipMemcpyDeviceToDevice, st_));
        HIPCALL(hipMemcpyAsync(f_.zzm, o.f_.zzm, (size_t)d_.ni * d_.nj * sizeof(double), hipMemcpyDeviceToDevice, st_));
    }
    HIPCALL(hipMemcpyAsync(sc_, o.sc_, sizeof(DevScalars), hipMemcpyDeviceToDevice, st_));
    // grid (moving mesh, powder layers) and host state
    p_.delt = o.p_.delt;
    p_.outtime = o.p_.outtime;
    std::memcpy(p_.xzone, o.p_.xzone, sizeof p_.xzone);
    std::memcpy(p_.yzone, o.p_.yzone, sizeof p_.yzone);
    hg_ = o.hg_;
    set_active_nk(o.d_.nk);
    int st = upload_geometry();
    if (st) return st;
    timet_ = o.timet_;
    scanvel_ = o.scanvel_;
    beamx_ = o.beamx_; beamy_ = o.beamy_; beamz_ = o.beamz_;
    oldx_ = o.oldx_; oldy_ = o.oldy_; oldz_ = o.oldz_;
    powerindicator_ = o.powerindicator_;
    laser_init_ = o.laser_init_;
    trackindx_ = o.trackindx_;
    istart_ = o.istart_;
    jstart_ = o.jstart_;
    heat_it_ = o.heat_it_;
    dt_state_ = o.dt_state_;
    step_ = o.step_;
    niter_ = o.niter_;
    itertot_ = o.itertot_;
    last_solve_rc_ = o.last_solve_rc_;
    alen_ = o.alen_;
    steadylength_ = o.steadylength_;
    tracksect_ = o.tracksect_;
    section_sci_ = o.section_sci_;
    box_ = o.box_;
    HIPCALL(hipStreamSynchronize(st_));
    *sc_host_ = *o.sc_host_;
    release_graph();
    return AMCFD_OK;
}

int Solver::raise_sweep_error() {
    const int one = 1;
    HIPCALL(hipMemcpyAsync(&sc_->error, &one, sizeof one, hipMemcpyHostToDevice, st_));
    HIPCALL(hipStreamSynchronize(st_));
    return AMCFD_OK;
}

int Solver::read_scalars() {
    HIPCALL(hipMemcpyAsync(sc_host_, sc_, sizeof(DevScalars), hipMemcpyDeviceToHost, st_));
    HIPCALL(hipStreamSynchronize(st_));
    if (sc_host_->error) return set_error(AMCFD_E_HIP, "line-sweep dependency wait expired");
    return 0;
}

// ---------------------------------------------------------------------------
// momentum + pressure correction on the melt-pool box, CFDSolverManager.cpp:660-675
// ---------------------------------------------------------------------------
int Solver::flow_iteration() {
    box_ = Box{sc_host_->istat, sc_host_->iend, sc_host_->jstat, sc_host_->jend, sc_host_->kstat};
    pr_.nblk_f = flow_partial_blocks(box_, d_);
    launch_clean_velocities(st_, d_, f_, box_, p_.tsolid);
    double *phis[5] = {nullptr, f_.uvel, f_.vvel, f_.wvel, f_.pp};
    static const char *names[5] = {"", "sweep_u", "sweep_v", "sweep_w", "sweep_p"};
    for (int ivar = 1; ivar <= 4; ++ivar) {
        launch_bc_flow(st_, d_, g_, f_, box_, p_, ivar);
        if (ivar < 4) launch_momentum(st_, d_, g_, f_, pr_, box_, p_, scanvel_, ivar);
        else launch_pressure(st_, d_, g_, f_, pr_, box_, p_);
        int st = sweep(phis[ivar], plan_box_, box_.istatp1(), box_.iendm1(), box_.jstat, box_.jend, box_.kstat,
                       d_.nkm2, 1, names[ivar]);
        if (st) return st;
        if (ivar == 4) launch_correct(st_, d_, f_, box_, p_);
    }
    HIPCALL(hipGetLastError());
    return 0;
}

// the learned warm start around a solve: record mode keeps the state the
// step starts from (and, after the solve, its increment) as a training
// sample; apply mode replaces the iterate the step starts from (the state
// the previous step ended with, CFDSolverManager.cpp:608-612) by the
// network's guess, and brings T, the liquid fraction and the properties in
// line with it before the first iteration
//
// The guard of mode 2: a solve that converges ends on the reference's test
// (CFDSolverManager.cpp:690-716), whatever it started from; one that runs to
// maxit ends where maxit iterations from its start lead, so there the guess
// would change the answer.  Mode 2 therefore starts a step from the guess
// only when the step before it converged, and keeps the plain start on runs
// whose steps end at maxit.
int Solver::warm_start_begin() {
    if (!ws_) return 0;
    if (ws_mode_ == 1) return ws_capture_input(ws_, *this);
    if (ws_mode_ == 2 && last_solve_rc_ != AMCFD_OK) {
        ++ws_guarded_;
        return 0;
    }
    ++ws_applied_;
    int st = ws_apply(ws_, *this);
    if (st) return st;
    launch_h2t(st_, d_, g_, f_, pr_, p_);
    launch_props(st_, d_, f_, p_);
    HIPCALL(hipGetLastError());
    return 0;
}

int Solver::warm_start_end() {
    if (!ws_ || ws_mode_ != 1) return 0;
    return ws_capture_target(ws_, *this);
}

int Solver::solve_step() {
    int st0 = warm_start_begin();
    if (st0) return st0;
    int rc0 = mode_ == AMCFD_MODE_MIXED ? solve_step_mixed() : solve_step_fp64();
    if (rc0 < 0) return rc0;
    last_solve_rc_ = rc0;
    st0 = warm_start_end();
    return st0 ? st0 : rc0;
}

int Solver::solve_step_fp64() {
    HIPCALL(hipEventRecord(step_ev_[0], st_));
    double tdma_before = timers.count("sweep_h") ? timers["sweep_h"].total_ms : 0.0;
    const double alenold = alen_;
    niter_ = 0;
    int rc = AMCFD_I_MAXIT;
    while (niter_ < p_.maxit) {
        ++niter_;
        ++itertot_;
        SourceParams sp = source_params();
        int slot;
        launch_bc_enthalpy(st_, d_, g_, f_, pr_, p_);
        tic("assemble_h", &slot);
        launch_energy_assemble(st_, d_, g_, f_, pr_, p_, sp);
        toc("assemble_h", slot);
        launch_block_correction(st_, d_, f_, pr_, f_.enthalpy, shared_device_);
        int st = sweep(f_.enthalpy, plan_dom_, 1, d_.nim2, 1, d_.njm2, 1, d_.nkm2, 0, "sweep_h");
        if (st) return st;
        // the properties (CFDSolverManager.cpp:624) in the H -> T launch: the
        // block correction and the sweep read and write none of their fields
        tic("h2t", &slot);
        launch_props_h2t(st_, d_, g_, f_, pr_, p_);
        toc("h2t", slot);
        launch_melt_pool(st_, d_, g_, f_, pr_, p_, istart_, jstart_, sc_);
        HIPCALL(hipGetLastError());
        if ((st = read_scalars())) return st;
        alen_ = sc_host_->alen;
        int flow = 0;
        if (sc_host_->tpeak > p_.tsolid && p_.navierstokes) {
            if ((st = flow_iteration())) return st;
            flow = 1;
        }
        if (p_.species) {   // CFDSolverManager.cpp:677-686
            launch_bc_species(st_, d_, g_, f_, p_);
            tic("assemble_c", &slot);
            launch_species_assemble(st_, d_, g_, f_, pr_, p_);
            toc("assemble_c", slot);
            launch_block_correction(st_, d_, f_, pr_, f_.conc, shared_device_);
            if ((st = sweep(f_.conc, plan_dom_, 1, d_.nim2, 1, d_.njm2, 1, d_.nkm2, 0, "sweep_c"))) return st;
        }
        launch_finalize(st_, d_, pr_, sc_, p_, flow, p_.species, p_.steady, sc_host_->width);
        HIPCALL(hipGetLastError());
        if ((st = read_scalars())) return st;
        if (profile_) collect_timers();
        const DevScalars &s = *sc_host_;
        // std::max over {resorm, resoru, resorv, resorw, resorc}: first largest wins
        double res[5] = {s.resorm, s.resoru, s.resorv, s.resorw, s.resorc};
        double amaxres = res[0];
        for (int r = 1; r < 5; ++r)
            if (amaxres < res[r]) amaxres = res[r];
        const double ratio = s.ratio, tol = p_.nonlintol;
        bool done;
        if (p_.steady) done = ratio <= 1.001 && ratio >= 0.999 && amaxres < tol;
        else if (powerindicator_ == 1) done = ratio <= 1.01 && ratio >= 0.99 && amaxres < tol;
        else done = s.resorh < 1e-6 && amaxres < tol;
        if (done) {
            rc = AMCFD_OK;
            break;
        }
    }
    if (rc == AMCFD_I_MAXIT && std::fabs(alen_ - alenold) < 1.0e-6) steadylength_ = true;
    HIPCALL(hipEventRecord(step_ev_[1], st_));
    HIPCALL(hipEventSynchronize(step_ev_[1]));
    float ms = 0.f;
    HIPCALL(hipEventElapsedTime(&ms, step_ev_[0], step_ev_[1]));
    last_total_ms = ms;
    last_tdma_ms = profile_ ? (timers["sweep_h"].total_ms - tdma_before) : -1.0;
    if (!std::isfinite(sc_host_->resorh))
        return set_error(AMCFD_E_NAN, "the enthalpy residual is not finite after step " + std::to_string(step_));
    return rc;
}

// ---------------------------------------------------------------------------
// mixed-precision mode: the same SIMPLE loop (CFDSolverManager.cpp:595-722)
// with FP32 enthalpy coefficients and delta-form line passes in parallel
// order (mixed.hip); properties, H->T, the melt pool, the heat balance, the
// residual folds and the convergence test are the FP64 mode's kernels
// ---------------------------------------------------------------------------
int Solver::energy_sweeps_mixed(const int *skip) {
    double *cur = f_.enthalpy, *other = mx_.halt;
    for (char c : mx_sched_) {
        if (c == 'x') {
            launch_mixed_xline(st_, d_, mx_, cur, 1, d_.nim2, 1, d_.njm2, 1, d_.nkm2, 0, skip, line_batch_);
            launch_mixed_xline(st_, d_, mx_, cur, 1, d_.nim2, 1, d_.njm2, 1, d_.nkm2, 1, skip, line_batch_);
        } else {
            launch_mixed_yzline(st_, d_, mx_, cur, other, c == 'y' ? 1 : 2, 1, d_.nim2, 1, d_.njm2, 1, d_.nkm2, skip);
            std::swap(cur, other);
        }
    }
    HIPCALL(hipGetLastError());
    return 0;
}

// One SIMPLE iteration of the mixed mode.  The melt-pool-box kernels always
// take the box from device memory (the melt-pool kernel writes it) with a
// fixed number of workgroups, so the host-driven and the captured loop run
// the very same launches.  Without ctl the host reads the scalars after the
// melt pool (flow or not) and after the finalize (convergence), as the FP64
// mode does; with ctl the kernels skip themselves on the device flags and
// the finalize launch takes the decision.
int Solver::enqueue_mixed_iteration(bool ctl, int it, bool ahead, bool fork_next) {
    const int *done = ctl ? &sc_->done : nullptr;
    int slot;
    if (!ctl) {
        SourceParams sp = source_params();
        *steph_ = StepDev{sp, powerindicator_, (int)p_.maxit, istart_, jstart_};
        HIPCALL(hipMemcpyAsync(stepd_, steph_, sizeof(StepDev), hipMemcpyHostToDevice, st_));
    }
    // the partials the finalize of this iteration folds (boundary and energy
    // residual) alternate with the next iteration's, which may be written
    // before this finalize runs
    Partials pq = pr_;
    if (it & 1) {
        pq.energy = energy_alt_;
        pq.bc = bc_alt_;
    }
    tic("assemble_h", &slot);
    if (!ahead) {
        launch_bc_enthalpy(st_, d_, g_, f_, pq, p_, done);
        launch_energy_assemble_mixed(st_, d_, g_, f_, pq, mx_, p_, SourceParams{}, done, stepd_);
    } else {   // the tiles near the box, once the flow part and the far tiles are done
        HIPCALL(hipStreamWaitEvent(st_, ev_join_, 0));
        launch_energy_assemble_mixed(st_, d_, g_, f_, pq, mx_, p_, SourceParams{}, done, stepd_, &sc_->istat, 2,
                                     ghost_save_, ghost_spec_);
    }
    toc("assemble_h", slot);
    launch_mixed_block_correction(st_, d_, mx_, f_.enthalpy, done);
    tic("sweep_h", &slot);
    int st = energy_sweeps_mixed(done);
    toc("sweep_h", slot);
    if (st) return st;
    // the properties from T (CFDSolverManager.cpp:624): nothing between the
    // assembly and here reads or writes what they read or write, so they run
    // after the line passes, in the H -> T launch
    tic("h2t", &slot);
    launch_props_h2t(st_, d_, g_, f_, pr_, p_, done, !mixed_has_volumetric_source(p_));
    toc("h2t", slot);
    launch_melt_pool(st_, d_, g_, f_, pr_, p_, istart_, jstart_, sc_, ctl, stepd_);
    HIPCALL(hipGetLastError());
    int flow = 1;
    if (!ctl) {
        if ((st = read_scalars())) return st;
        alen_ = sc_host_->alen;
        flow = sc_host_->tpeak > p_.tsolid && p_.navierstokes;
    }
    pr_.nblk_f = box_blocks_;
    pq.nblk_f = box_blocks_;
    if (fork_next) {
        // The next iteration's boundary launch and the tiles of its assembly
        // away from the box, alongside this iteration's flow part: they read
        // nothing the flow part writes, and write nothing it or this
        // iteration's finalize reads (its partials are the other set, its
        // source term is the energy equation's own).  The boundary launch
        // keeps the ghost values it overwrites, which the next iteration's
        // late assembly launch puts back if this iteration ends the loop.
        Partials pn = pr_;
        if (!(it & 1)) {
            pn.energy = energy_alt_;
            pn.bc = bc_alt_;
        }
        HIPCALL(hipEventRecord(ev_fork_, st_));
        HIPCALL(hipStreamWaitEvent(side_st_, ev_fork_, 0));
        launch_bc_enthalpy(side_st_, d_, g_, f_, pn, p_, done, ghost_save_, ghost_spec_);
        launch_energy_assemble_mixed(side_st_, d_, g_, f_, pn, mx_, p_, SourceParams{}, done, stepd_, &sc_->istat, 1,
                                     nullptr, nullptr, fork_blocks_);
        HIPCALL(hipEventRecord(ev_join_, side_st_));
    }
    if (flow) {
        BoxLaunch bl;
        bl.dbox = &sc_->istat;
        bl.skip = ctl ? &sc_->noflow : nullptr;
        bl.nblocks = box_blocks_;
        launch_clean_velocities(st_, d_, f_, box_, p_.tsolid, bl);
        double *phis[5] = {nullptr, f_.uvel, f_.vvel, f_.wvel, f_.pp};
        BoxPrep prep;
        if (box_fused_) {   // the boundary launch of each variable prepares its box passes
            if ((st = ensure_box_buffers(mx_box_passes_))) return st;
            prep = BoxPrep{box_bufs_, (size_t)d_.ni * d_.nj * d_.nkf, mx_box_passes_, box_claim_};
        }
        for (int ivar = 1; ivar <= 4; ++ivar) {
            launch_bc_flow(st_, d_, g_, f_, box_, p_, ivar, bl, box_fused_ ? &prep : nullptr);
            if (ivar < 4) launch_momentum(st_, d_, g_, f_, pr_, box_, p_, scanvel_, ivar, bl);
            else launch_pressure(st_, d_, g_, f_, pr_, box_, p_, bl);
            if (box_fused_) {
                if (launch_box_passes(st_, d_, f_, phis[ivar], prep.bufs, prep.bstride, prep.passes, prep.claim,
                                      &sc_->error, bl.skip, bl.dbox, box_blocks_, false))
                    return set_error(AMCFD_E_HIP, "box passes: launch failed");
            } else {
                for (int pass = 0; pass < mx_box_passes_; ++pass)
                    for (int colour = 0; colour < 2; ++colour)
                        launch_box_xline(st_, d_, f_, phis[ivar], 1, d_.nim2, 1, d_.njm2, 1, d_.nkm2, colour,
                                         bl.skip, bl.dbox, box_blocks_);
            }
            if (ivar == 4) launch_correct(st_, d_, f_, box_, p_, bl);
        }
    }
    launch_finalize(st_, d_, pq, sc_, p_, flow, 0, p_.steady, ctl ? 0.0 : sc_host_->width, ctl,
                    ctl ? stepd_ : nullptr);
    HIPCALL(hipGetLastError());
    return 0;
}

int Solver::solve_step_mixed() {
    // captured loop where the step's inputs are fixed for the whole step (a
    // heat-source table may advance its row between iterations, which the
    // host loop does as the reference)
    if (use_graph_ && heatt_.empty()) return solve_step_mixed_graph();
    HIPCALL(hipEventRecord(step_ev_[0], st_));
    const double alenold = alen_;
    niter_ = 0;
    int rc = AMCFD_I_MAXIT;
    while (niter_ < p_.maxit) {
        ++niter_;
        ++itertot_;
        int st = enqueue_mixed_iteration(false);
        if (st) return st;
        if ((st = read_scalars())) return st;
        if (profile_) collect_timers();
        const DevScalars &s = *sc_host_;
        double res[4] = {s.resorm, s.resoru, s.resorv, s.resorw};
        double amaxres = res[0];
        for (int r = 1; r < 4; ++r)
            if (amaxres < res[r]) amaxres = res[r];
        const double ratio = s.ratio, tol = p_.nonlintol;
        bool done;
        if (p_.steady) done = ratio <= 1.001 && ratio >= 0.999 && amaxres < tol;
        else if (powerindicator_ == 1) done = ratio <= 1.01 && ratio >= 0.99 && amaxres < tol;
        else done = s.resorh < 1e-6 && amaxres < tol;
        if (done) {
            rc = AMCFD_OK;
            break;
        }
    }
    if (rc == AMCFD_I_MAXIT && std::fabs(alen_ - alenold) < 1.0e-6) steadylength_ = true;
    HIPCALL(hipEventRecord(step_ev_[1], st_));
    HIPCALL(hipEventSynchronize(step_ev_[1]));
    float ms = 0.f;
    HIPCALL(hipEventElapsedTime(&ms, step_ev_[0], step_ev_[1]));
    last_total_ms = ms;
    last_tdma_ms = -1.0;
    if (!std::isfinite(sc_host_->resorh))
        return set_error(AMCFD_E_NAN, "the enthalpy residual is not finite after step " + std::to_string(step_));
    return rc;
}

int Solver::capture_mixed_graph() {
    release_graph();
    const int prof = profile_;
    profile_ = 0;   // no events inside a capture
    HIPCALL(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    int st = 0;
    // each iteration but the block's last runs the next one's boundary
    // launch and far assembly tiles alongside its own flow part
    const bool fork = fork_key() != 0;
    for (int it = 0; it < graph_block_ && !st; ++it)
        st = enqueue_mixed_iteration(true, it, it > 0 && fork, it + 1 < graph_block_ && fork);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st_, &g);
    profile_ = prof;
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return set_error(AMCFD_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    graph_ = g;
    HIPCALL(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
    gkey_.nk = d_.nk;
    gkey_.blocks = graph_block_;
    gkey_.box_blocks = box_blocks_;
    gkey_.delt = p_.delt;
    gkey_.scanvel = scanvel_;
    gkey_.sched = mx_sched_;
    gkey_.box_passes = mx_box_passes_;
    gkey_.box_fused = box_fused_;
    gkey_.line_batch = line_batch_;
    gkey_.fork_asm = fork_key();
    return 0;
}

int Solver::solve_step_mixed_graph() {
    HIPCALL(hipEventRecord(step_ev_[0], st_));
    const double alenold = alen_;
    // every launch argument the graph holds by value: rebuild when one moved
    if (!gexec_ || gkey_.nk != d_.nk || gkey_.blocks != graph_block_ || gkey_.box_blocks != box_blocks_ ||
        gkey_.delt != p_.delt || gkey_.scanvel != scanvel_ || gkey_.sched != mx_sched_ ||
        gkey_.box_passes != mx_box_passes_ || gkey_.box_fused != box_fused_ || gkey_.line_batch != line_batch_ ||
        gkey_.fork_asm != fork_key()) {
        int st = capture_mixed_graph();
        if (st) return st;
    }
    SourceParams sp = source_params();
    *steph_ = StepDev{sp, powerindicator_, (int)p_.maxit, istart_, jstart_};
    HIPCALL(hipMemcpyAsync(stepd_, steph_, sizeof(StepDev), hipMemcpyHostToDevice, st_));
    launch_loop_reset(st_, sc_);
    HIPCALL(hipGetLastError());
    const int nlaunch = ((int)p_.maxit + graph_block_ - 1) / graph_block_;
    for (int l = 0; l < nlaunch; ++l) {
        HIPCALL(hipGraphLaunch(gexec_, st_));
        int st = read_scalars();
        if (st) return st;
        if (sc_host_->done) break;
    }
    niter_ = sc_host_->niter;
    itertot_ += niter_;
    alen_ = sc_host_->alen;
    const int rc = sc_host_->converged ? AMCFD_OK : AMCFD_I_MAXIT;
    if (rc == AMCFD_I_MAXIT && std::fabs(alen_ - alenold) < 1.0e-6) steadylength_ = true;
    HIPCALL(hipEventRecord(step_ev_[1], st_));
    HIPCALL(hipEventSynchronize(step_ev_[1]));
    float ms = 0.f;
    HIPCALL(hipEventElapsedTime(&ms, step_ev_[0], step_ev_[1]));
    last_total_ms = ms;
    last_tdma_ms = -1.0;
    if (!std::isfinite(sc_host_->resorh))
        return set_error(AMCFD_E_NAN, "the enthalpy residual is not finite after step " + std::to_string(step_));
    return rc;
}

// ---------------------------------------------------------------------------
// grouped runs: the captured loop's pieces for a runner that launches each
// kernel of the iteration once for many solvers (group.cpp)
// ---------------------------------------------------------------------------
int Solver::group_check() const {
    if (mode_ != AMCFD_MODE_MIXED) return set_error(AMCFD_E_UNSUPPORTED, "grouped runs take mixed-mode solvers");
    if (!use_graph_) return set_error(AMCFD_E_UNSUPPORTED, "grouped runs take solvers with the captured loop on");
    if (!heatt_.empty())
        return set_error(AMCFD_E_UNSUPPORTED, "a heat-source table advances between iterations (host loop)");
    if (mx_sched_.find_first_of("yz") != std::string::npos)
        return set_error(AMCFD_E_UNSUPPORTED, "y/z line passes have no grouped form");
    return 0;
}

std::string Solver::group_key() const {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%p:%d:%d:%d:%a:%a:%s:%d:%d:%d:%p;", (const void *)this, d_.nk, graph_block_,
                  box_blocks_, p_.delt, scanvel_, mx_sched_.c_str(), mx_box_passes_, box_fused_, line_batch_,
                  (const void *)box_bufs_);
    return buf;
}

int Solver::group_begin_solve(hipEvent_t ready) {
    int st = warm_start_begin();
    if (st) return st;
    group_alenold_ = alen_;
    SourceParams sp = source_params();
    *steph_ = StepDev{sp, powerindicator_, (int)p_.maxit, istart_, jstart_};
    HIPCALL(hipMemcpyAsync(stepd_, steph_, sizeof(StepDev), hipMemcpyHostToDevice, st_));
    launch_loop_reset(st_, sc_);
    HIPCALL(hipGetLastError());
    HIPCALL(hipEventRecord(ready, st_));
    return 0;
}

int Solver::group_record_iteration(LaunchRecorder &rec) {
    if (int st = ensure_box_buffers(mx_box_passes_)) return st;   // allocated before any recording
    const int prof = profile_;
    profile_ = 0;
    tls_recorder() = &rec;
    const int st = enqueue_mixed_iteration(true);
    tls_recorder() = nullptr;
    profile_ = prof;
    if (!st && rec.unsupported) return set_error(AMCFD_E_UNSUPPORTED, "an iteration launch has no grouped form");
    return st;
}

int Solver::group_end_solve(hipEvent_t done, float ms) {
    HIPCALL(hipStreamWaitEvent(st_, done, 0));
    int st = read_scalars();
    if (st) return st;
    niter_ = sc_host_->niter;
    itertot_ += niter_;
    alen_ = sc_host_->alen;
    const int rc = sc_host_->converged ? AMCFD_OK : AMCFD_I_MAXIT;
    if (rc == AMCFD_I_MAXIT && std::fabs(alen_ - group_alenold_) < 1.0e-6) steadylength_ = true;
    last_total_ms = ms;
    last_tdma_ms = -1.0;
    if (!std::isfinite(sc_host_->resorh))
        return set_error(AMCFD_E_NAN, "the enthalpy residual is not finite after step " + std::to_string(step_));
    last_solve_rc_ = rc;
    st = warm_start_end();
    return st ? st : rc;
}

int Solver::end_step() {
    // getSolidificationParameters runs between the solve and the push-back
    // (main.cpp:151-152), steady runs included; skipped while nothing melts
    if (p_.solidification && sc_host_->tpeak > p_.tsolid) {
        launch_solidification(st_, d_, g_, f_, p_.tsolid, scanvel_);
        HIPCALL(hipGetLastError());
    }
    if (p_.steady) {
        HIPCALL(hipStreamSynchronize(st_));
        return AMCFD_OK;
    }
    double trackid = 0.0;
    for (int r = p_.n_toolpath; r > 0; --r)
        if (timet_ <= tool_[6 * (r - 1)]) trackid = (double)(int)tool_[6 * (r - 1) + 5];
    launch_end_step(st_, d_, f_, p_, (double)trackindx_, trackid);
    if (p_.species) launch_species_end(st_, d_, f_, p_);
    HIPCALL(hipGetLastError());
    // the step is complete when this returns, as the reference's is
    HIPCALL(hipStreamSynchronize(st_));
    return AMCFD_OK;
}

// ---------------------------------------------------------------------------
// field access
// ---------------------------------------------------------------------------
double *Solver::field_ptr(const char *n) {
    struct {
        const char *name;
        double *p;
    } t[] = {{"enthalpy", f_.enthalpy}, {"temp", f_.temp}, {"fracl", f_.fracl}, {"uvel", f_.uvel},
             {"vvel", f_.vvel}, {"wvel", f_.wvel}, {"pressure", f_.pressure}, {"pp", f_.pp}, {"rho", f_.rho},
             {"diff", f_.diff}, {"vis", f_.vis}, {"hnot", f_.hnot}, {"fraclnot", f_.fraclnot},
             {"unot", f_.unot}, {"vnot", f_.vnot}, {"wnot", f_.wnot}, {"sourceinput", f_.sourceinput},
             {"dux", f_.dux}, {"dvy", f_.dvy}, {"dwz", f_.dwz}, {"csfrac", f_.csfrac}, {"solfrac", f_.solfrac},
             {"mpPhase", f_.mpphase}, {"mpLiqPhase", f_.mpliqphase}, {"ap", f_.ap}, {"an", f_.an},
             {"as", f_.as}, {"ae", f_.ae}, {"aw", f_.aw}, {"at", f_.at}, {"ab", f_.ab}, {"su", f_.su},
             {"avgconcentration", f_.conc}, {"avgconcentrationnot", f_.concnot}, {"concentration", f_.concl},
             {"grad", f_.grad}, {"rate", f_.rate}, {"tempr", f_.tempr}, {"zr", f_.zr}};
    // species and G/R fields exist only with their keywords (null otherwise)
    for (auto &e : t)
        if (std::strcmp(e.name, n) == 0) return e.p;
    return nullptr;
}

int Solver::get_field(const char *name, double *host, size_t n) {
    size_t nn = (size_t)d_.ni * d_.nj * d_.nkf;
    const bool fs_name = !std::strcmp(name, "tempr") || !std::strcmp(name, "zr");
    if (fs_name && !p_.energyfreesurface) {
        // registered whatever the deck (FreeSurfaceManager.cpp:35-67) and never
        // updated without the free surface: the initial temperature and z
        if (n != nn) return set_error(AMCFD_E_ARG, "field buffer must hold ni*nj*nk values");
        for (size_t c = 0; c < nn; ++c)
            host[c] = name[0] == 'z' ? hg0_z_[c / ((size_t)d_.ni * d_.nj)] : p_.temppreheat;
        return AMCFD_OK;
    }
    double *p = field_ptr(name);
    if (!p) return set_error(AMCFD_E_NAME, std::string("unknown field ") + name);
    if (n != nn) return set_error(AMCFD_E_ARG, "field buffer must hold ni*nj*nk values");
    HIPCALL(hipMemcpyAsync(host, p, nn * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCALL(hipStreamSynchronize(st_));
    return AMCFD_OK;
}

int Solver::set_field(const char *name, const double *host, size_t n) {
    double *p = field_ptr(name);
    if (!p) return set_error(AMCFD_E_NAME, std::string("unknown field ") + name);
    size_t nn = (size_t)d_.ni * d_.nj * d_.nkf;
    if (n != nn) return set_error(AMCFD_E_ARG, "field buffer must hold ni*nj*nk values");
    HIPCALL(hipMemcpyAsync(p, host, nn * sizeof(double), hipMemcpyHostToDevice, st_));
    HIPCALL(hipStreamSynchronize(st_));
    return AMCFD_OK;
}

int Solver::get_geometry(const char *name, double *host, size_t n) {
    struct {
        const char *name;
        const std::vector<double> *v;
        size_t len;
    } t[] = {{"x", &hg_.x, (size_t)d_.ni}, {"y", &hg_.y, (size_t)d_.nj}, {"z", &hg_.z, (size_t)d_.nkf},
             {"xu", &hg_.xu, (size_t)d_.ni}, {"yv", &hg_.yv, (size_t)d_.nj}, {"zw", &hg_.zw, (size_t)d_.nkf},
             {"dxpwinv", &hg_.dxpwinv, (size_t)d_.ni}, {"dypsinv", &hg_.dypsinv, (size_t)d_.nj},
             {"dzpbinv", &hg_.dzpbinv, (size_t)d_.nkf}, {"fracx", &hg_.fracx, (size_t)d_.ni},
             {"fracy", &hg_.fracy, (size_t)d_.nj}, {"fracz", &hg_.fracz, (size_t)d_.nkf},
             {"areaij", &hg_.areaij, hg_.areaij.size()}, {"areajk", &hg_.areajk, hg_.areajk.size()},
             {"areaik", &hg_.areaik, hg_.areaik.size()}};
    for (auto &e : t)
        if (std::strcmp(e.name, name) == 0) {
            if (n != e.len) return set_error(AMCFD_E_ARG, "geometry buffer has the wrong length");
            std::copy(e.v->begin(), e.v->begin() + e.len, host);
            return AMCFD_OK;
        }
    if (std::strcmp(name, "volume") == 0) {
        size_t nn = (size_t)d_.ni * d_.nj * d_.nkf;
        if (n != nn) return set_error(AMCFD_E_ARG, "geometry buffer has the wrong length");
        std::fill(host, host + nn, 0.0);
        for (int k = 1; k < d_.nkf - 1; ++k)
            for (int j = 1; j < d_.nj - 1; ++j)
                for (int i = 1; i < d_.ni - 1; ++i) host[((size_t)k * d_.nj + j) * d_.ni + i] = hg_.volume(i, j, k);
        return AMCFD_OK;
    }
    return set_error(AMCFD_E_NAME, std::string("unknown geometry array ") + name);
}

int Solver::get_scalars(amcfd_scalars *o) {
    HIPCALL(hipMemcpyAsync(sc_host_, sc_, sizeof(DevScalars), hipMemcpyDeviceToHost, st_));
    HIPCALL(hipStreamSynchronize(st_));
    const DevScalars &s = *sc_host_;
    std::memset(o, 0, sizeof(*o));
    o->timet = timet_;
    o->delt = p_.delt;
    o->outtime = p_.outtime;
    o->niter = niter_;
    o->itertot = itertot_;
    o->resorh = s.resorh;
    o->resorm = s.resorm;
    o->resoru = s.resoru;
    o->resorv = s.resorv;
    o->resorw = s.resorw;
    o->resorc = s.resorc;
    o->tpeak = s.tpeak;
    o->alen = s.alen;
    o->depth = s.depth;
    o->width = s.width;
    o->ratio = s.ratio;
    o->heatinlaser = s.heatinlaser;
    o->heatout = s.heatout;
    o->accul = s.accul;
    o->heatvol = s.heatvol;
    o->ahtoploss = s.ahtoploss;
    o->beamposx = beamx_;
    o->beamposy = beamy_;
    o->scanvel = scanvel_;
    o->powerindicator = powerindicator_;
    o->trackindx = trackindx_;
    o->istat = s.istat;
    o->iend = s.iend;
    o->jstat = s.jstat;
    o->jend = s.jend;
    o->kstat = s.kstat;
    o->istart = istart_;
    o->jstart = jstart_;
    o->nk = d_.nk;
    o->step = step_;
    return AMCFD_OK;
}

}  // namespace amcfd
