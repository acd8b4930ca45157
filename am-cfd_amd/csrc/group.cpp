// Grouped runs: many independent mixed-mode solvers (instances of a
// parameter sweep, uneven grids) advance in lock step, and every kernel of a
// SIMPLE iteration is one launch over all of them.
//
// The runner has each solver record one iteration of its captured loop (the
// launches it would issue, group.h) and lines the records up by position:
// position p of every solver is the same kernel family, so one grouped
// launch of that family with one argument entry per solver replaces n
// launches (solvers whose decks choose another kernel at a position, such as
// another discretisation scheme's assembly, get one launch per kernel
// there).  Where solvers differ in a kernel's size class (the row capacity
// of the line solvers, which follows the grid's x extent), the largest class
// serves all: the line solvers take the rows per lane from the line itself,
// so a larger capacity computes the same bits.  A graph of
// graph_block such iterations plus a kernel that gathers the solvers' done
// flags into host memory is captured once and relaunched until every
// solver's step is done.  Each solver computes exactly what it computes
// alone: the same kernels with the same arguments in the same order, and
// results that do not depend on which workgroup serves which tile.
//
// Per step: every solver's host-side step start and device inputs on its own
// stream (laser update, warm start, StepDev, loop reset), the group stream
// waits for all of them, runs the graph, and the solvers' streams wait for
// the group before reading their scalars and pushing the step back.
#include "host.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace amcfd {

LaunchRecorder *&tls_recorder() {
    static thread_local LaunchRecorder *r = nullptr;
    return r;
}

namespace {

#define GCALL(x)                                                                                               \
    do {                                                                                                       \
        hipError_t e_ = (x);                                                                                   \
        if (e_ != hipSuccess) return set_error(AMCFD_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// one grouped launch: a kernel family at one position of the iteration and
// the instances whose launch there is of that family
struct Launch {
    const void *kernel = nullptr;
    size_t pos = 0;
    int family = 0, rank = -1;
    dim3 block;
    size_t lds = 0, entry_bytes = 0, offset = 0;
    long long blocks = 0;
    std::vector<int> inst;
};

class GroupRun {
  public:
    ~GroupRun() { release(); }
    int run(const std::vector<Solver *> &s, int max_steps, int *steps_done);

  private:
    hipStream_t st_ = nullptr;
    hipEvent_t t0_ = nullptr, t1_ = nullptr;
    std::vector<hipEvent_t> ready_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    std::string key_;
    void *entries_ = nullptr;
    int **done_ptrs_ = nullptr;
    int *all_done_ = nullptr;   // pinned host memory the done gather writes
    int iters_ = 8;

    int setup();
    int build(const std::vector<Solver *> &act);
    void release_graph();
    void release();
};

void GroupRun::release_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    gexec_ = nullptr;
    graph_ = nullptr;
    if (entries_) (void)hipFree(entries_);
    if (done_ptrs_) (void)hipFree(done_ptrs_);
    entries_ = nullptr;
    done_ptrs_ = nullptr;
    key_.clear();
}

void GroupRun::release() {
    if (st_) (void)hipStreamSynchronize(st_);
    release_graph();
    for (hipEvent_t e : ready_) (void)hipEventDestroy(e);
    ready_.clear();
    if (t0_) (void)hipEventDestroy(t0_);
    if (t1_) (void)hipEventDestroy(t1_);
    if (all_done_) (void)hipHostFree(all_done_);
    if (st_) (void)hipStreamDestroy(st_);
    t0_ = t1_ = nullptr;
    all_done_ = nullptr;
    st_ = nullptr;
}

int GroupRun::setup() {
    if (st_) return 0;
    GCALL(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    GCALL(hipEventCreate(&t0_));
    GCALL(hipEventCreate(&t1_));
    GCALL(hipHostMalloc((void **)&all_done_, sizeof(int)));
    *all_done_ = 0;
    return 0;
}

// record one iteration of every solver, merge by position, upload the
// argument entries and capture iters_ iterations of grouped launches
int GroupRun::build(const std::vector<Solver *> &act) {
    release_graph();
    const int n = (int)act.size();
    std::vector<LaunchRecorder> rec(n);
    for (int i = 0; i < n; ++i)
        if (int st = act[i]->group_record_iteration(rec[i])) return st;
    const size_t npos = rec[0].recs.size();
    for (int i = 1; i < n; ++i)
        if (rec[i].recs.size() != npos)
            return set_error(AMCFD_E_UNSUPPORTED, "grouped solvers issue different launch sequences");
    // per position, one launch per kernel family among the instances (decks
    // with another discretisation scheme run another assembly kernel); the
    // largest size class of a family serves its instances
    std::vector<Launch> launches;
    size_t bytes = 0;
    for (size_t p = 0; p < npos; ++p) {
        const size_t first = launches.size();
        for (int i = 0; i < n; ++i) {
            const LaunchRecord &r = rec[i].recs[p];
            size_t l = first;
            while (l < launches.size() && launches[l].family != r.family) ++l;
            if (l == launches.size()) {
                launches.emplace_back();
                launches[l].pos = p;
                launches[l].family = r.family;
                launches[l].block = r.block;
                launches[l].lds = r.lds;
                launches[l].entry_bytes = r.entry.size();
            }
            Launch &q = launches[l];
            if (r.block.x != q.block.x || r.block.y != q.block.y || r.block.z != q.block.z || r.lds != q.lds ||
                r.entry.size() != q.entry_bytes)
                return set_error(AMCFD_E_UNSUPPORTED, "grouped solvers launch one kernel family in different shapes");
            if (r.rank > q.rank) {
                q.rank = r.rank;
                q.kernel = r.group_kernel;
            }
            q.blocks += (long long)r.grid.x * r.grid.y * r.grid.z;
            q.inst.push_back(i);
        }
        for (size_t l = first; l < launches.size(); ++l) {
            if (launches[l].blocks > 0x7fffffffLL)
                return set_error(AMCFD_E_CAPACITY, "grouped launch beyond the grid limit");
            launches[l].offset = bytes;
            bytes += ((launches[l].entry_bytes * launches[l].inst.size() + 255) / 256) * 256;
        }
    }
    // every entry starts with its instance's first block (int)
    std::vector<unsigned char> host(bytes, 0);
    for (const Launch &q : launches) {
        int start = 0;
        for (size_t e = 0; e < q.inst.size(); ++e) {
            const LaunchRecord &r = rec[q.inst[e]].recs[q.pos];
            unsigned char *dst = host.data() + q.offset + e * q.entry_bytes;
            std::memcpy(dst, r.entry.data(), r.entry.size());
            std::memcpy(dst, &start, sizeof(int));
            start += (int)(r.grid.x * r.grid.y * r.grid.z);
        }
    }
    GCALL(hipMalloc(&entries_, std::max<size_t>(bytes, 256)));
    GCALL(hipMemcpy(entries_, host.data(), bytes, hipMemcpyHostToDevice));
    std::vector<int *> done(n);
    for (int i = 0; i < n; ++i) done[i] = act[i]->group_done_flag();
    GCALL(hipMalloc((void **)&done_ptrs_, sizeof(int *) * n));
    GCALL(hipMemcpy(done_ptrs_, done.data(), sizeof(int *) * n, hipMemcpyHostToDevice));

    GCALL(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int it = 0; it < iters_ && e == hipSuccess; ++it)
        for (size_t l = 0; l < launches.size() && e == hipSuccess; ++l) {
            const Launch &q = launches[l];
            if (q.blocks == 0) continue;
            void *ptr = (unsigned char *)entries_ + q.offset;
            int cnt = (int)q.inst.size();
            void *args[] = {&ptr, &cnt};
            e = hipLaunchKernel(q.kernel, dim3((unsigned)q.blocks), q.block, args, q.lds, st_);
        }
    if (e == hipSuccess) {
        launch_group_done(st_, done_ptrs_, n, all_done_);
        e = hipGetLastError();
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st_, &g);
    if (e != hipSuccess || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return set_error(AMCFD_E_HIP, std::string("grouped capture: ") + hipGetErrorString(e != hipSuccess ? e : ec));
    }
    graph_ = g;
    GCALL(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
    return 0;
}

int GroupRun::run(const std::vector<Solver *> &s, int max_steps, int *steps_done) {
    if (int st = setup()) return st;
    const int n = (int)s.size();
    for (int i = 0; i < n; ++i) steps_done[i] = 0;
    for (;;) {
        std::vector<Solver *> act;
        std::vector<int> idx;
        for (int i = 0; i < n; ++i)
            if ((max_steps < 0 || steps_done[i] < max_steps) && !s[i]->finished()) {
                act.push_back(s[i]);
                idx.push_back(i);
            }
        if (act.empty()) return AMCFD_OK;
        const int na = (int)act.size();
        while ((int)ready_.size() < na) {
            hipEvent_t ev;
            GCALL(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            ready_.push_back(ev);
        }
        int maxit = 0, iters = 1;
        std::string key;
        for (int a = 0; a < na; ++a) {
            Solver &v = *act[a];
            int st = v.begin_step();
            if (st < 0) return set_error(st, "grouped instance " + std::to_string(idx[a]) + ": " + amcfd_last_error());
            if ((st = v.group_begin_solve(ready_[a])))
                return set_error(st, "grouped instance " + std::to_string(idx[a]) + ": " + amcfd_last_error());
            GCALL(hipStreamWaitEvent(st_, ready_[a], 0));
            int on, block, bb;
            v.graph_settings(&on, &block, &bb);
            iters = std::max(iters, block);
            maxit = std::max(maxit, (int)v.params().maxit);
            key += v.group_key();
        }
        key += std::to_string(iters);
        if (key != key_) {
            iters_ = iters;
            if (int st = build(act)) return st;
            key_ = key;
        }
        GCALL(hipEventRecord(t0_, st_));
        *all_done_ = 0;
        const int nlaunch = (maxit + iters_ - 1) / iters_;
        for (int l = 0; l < nlaunch; ++l) {
            GCALL(hipGraphLaunch(gexec_, st_));
            GCALL(hipStreamSynchronize(st_));
            if (*all_done_) break;
        }
        GCALL(hipEventRecord(t1_, st_));
        GCALL(hipEventSynchronize(t1_));
        float ms = 0.f;
        GCALL(hipEventElapsedTime(&ms, t0_, t1_));
        for (int a = 0; a < na; ++a) {
            int rc = act[a]->group_end_solve(t1_, ms);
            if (rc < 0) return set_error(rc, "grouped instance " + std::to_string(idx[a]) + ": " + amcfd_last_error());
            if ((rc = act[a]->end_step()) < 0)
                return set_error(rc, "grouped instance " + std::to_string(idx[a]) + ": " + amcfd_last_error());
            ++steps_done[idx[a]];
        }
    }
}

}  // namespace

int run_grouped(const std::vector<Solver *> &s, int max_steps, int *steps_done) {
    for (size_t i = 0; i < s.size(); ++i)
        if (int st = s[i]->group_check())
            return set_error(st, "grouped instance " + std::to_string(i) + ": " + amcfd_last_error());
    if (s.empty()) return AMCFD_OK;
    GroupRun g;
    return g.run(s, max_steps, steps_done);
}

}  // namespace amcfd
