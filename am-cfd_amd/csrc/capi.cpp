// extern "C" boundary of the library (declared in include/amcfd.h).

#include "host.h"

struct amcfd_ws;
namespace amcfd {
WarmStart *ws_of(amcfd_ws *h);
}

#include <cstring>
#include <exception>
#include <new>
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace amcfd {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

}  // namespace amcfd

using amcfd::Deck;
using amcfd::Solver;

struct amcfd_deck {
    Deck d;
};
struct amcfd_solver {
    Solver s;
};

#define GUARD(expr)                                                              \
    try {                                                                        \
        return (expr);                                                           \
    } catch (const std::bad_alloc &) {                                           \
        return amcfd::set_error(AMCFD_E_CAPACITY, "host allocation failed");     \
    } catch (const std::exception &e) {                                          \
        return amcfd::set_error(AMCFD_E_PARSE, e.what());                        \
    }

extern "C" {

int amcfd_deck_load(const char *deck_path, amcfd_deck **out) {
    if (!deck_path || !out) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    *out = nullptr;
    try {
        amcfd_deck *d = new amcfd_deck();
        int st = d->d.load(deck_path);
        if (st) {
            delete d;
            return st;
        }
        *out = d;
        return AMCFD_OK;
    } catch (const std::bad_alloc &) {
        return amcfd::set_error(AMCFD_E_CAPACITY, "host allocation failed");
    } catch (const std::exception &e) {
        return amcfd::set_error(AMCFD_E_PARSE, std::string("deck parse error: ") + e.what());
    }
}

const amcfd_params *amcfd_deck_params(const amcfd_deck *deck) { return deck ? &deck->d.p : nullptr; }

void amcfd_deck_free(amcfd_deck *deck) { delete deck; }

int amcfd_create(const amcfd_params *params, int mode, amcfd_solver **out) {
    if (!params || !out) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    *out = nullptr;
    try {
        amcfd_solver *s = new amcfd_solver();
        int st = s->s.create(*params, mode);
        if (st) {
            delete s;
            return st;
        }
        *out = s;
        return AMCFD_OK;
    } catch (const std::bad_alloc &) {
        return amcfd::set_error(AMCFD_E_CAPACITY, "host allocation failed");
    }
}

void amcfd_destroy(amcfd_solver *s) { delete s; }

int amcfd_begin_step(amcfd_solver *s) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    GUARD(s->s.begin_step());
}

int amcfd_solve_step(amcfd_solver *s) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    GUARD(s->s.solve_step());
}

int amcfd_end_step(amcfd_solver *s) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    GUARD(s->s.end_step());
}

int amcfd_step(amcfd_solver *s, int *finished) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (s->s.finished()) {
        if (finished) *finished = 1;
        return AMCFD_OK;
    }
    int st = amcfd_begin_step(s);
    if (st < 0) return st;
    int rs = amcfd_solve_step(s);
    if (rs < 0) return rs;
    st = amcfd_end_step(s);
    if (st < 0) return st;
    if (finished) *finished = s->s.finished();
    return rs;
}

int amcfd_run_batch(amcfd_solver *const *s, int n, int max_steps, int workers, int *steps_done) {
    if (n < 0 || (n > 0 && (!s || !steps_done))) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    for (int i = 0; i < n; ++i)
        if (!s[i]) return amcfd::set_error(AMCFD_E_ARG, "null solver in batch");
    {   // two workers must never step one solver (and its stream) at once
        std::vector<const amcfd_solver *> seen(s, s + n);
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            return amcfd::set_error(AMCFD_E_ARG, "the same solver appears twice in the batch");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return amcfd::set_error(AMCFD_E_HIP, "hipGetDevice failed");
    if (workers <= 0) workers = 16;
    workers = std::min(workers, n);
    // largest grids first, so that the last instances to start are short
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        int ia, ja, ka, ib, jb, kb;
        s[a]->s.dims(&ia, &ja, &ka);
        s[b]->s.dims(&ib, &jb, &kb);
        return (long long)ia * ja * ka > (long long)ib * jb * kb;
    });
    std::vector<int> rc(n, AMCFD_OK);
    std::vector<std::string> msg(n);
    std::atomic<int> next{0};
    auto worker = [&]() {
        // the calling thread's device, for the allocations a plan change makes
        const bool dev_ok = hipSetDevice(dev) == hipSuccess;
        for (int q = next++; q < n; q = next++) {
            const int i = order[q];
            if (!dev_ok) {
                rc[i] = AMCFD_E_HIP;
                msg[i] = "hipSetDevice failed";
                steps_done[i] = 0;
                continue;
            }
            int k = 0, fin = 0;
            while ((max_steps < 0 || k < max_steps) && !s[i]->s.finished()) {
                const int st = amcfd_step(s[i], &fin);
                if (st < 0) {
                    rc[i] = st;
                    msg[i] = amcfd_last_error();
                    break;
                }
                ++k;
            }
            steps_done[i] = k;
        }
    };
    for (int i = 0; i < n; ++i) s[i]->s.set_shared_device(workers > 1);
    // the pool outlives the try block, so threads already started are joined
    // before an error returns (a joinable std::thread must not be destroyed)
    std::vector<std::thread> pool;
    std::string start_error;
    try {
        pool.reserve(workers);
        for (int w = 0; w < workers; ++w) pool.emplace_back(worker);
    } catch (const std::exception &e) {
        start_error = e.what();
        next = n;   // no further instance is taken up
    }
    for (auto &t : pool) t.join();
    for (int i = 0; i < n; ++i) s[i]->s.set_shared_device(false);
    if (!start_error.empty())
        return amcfd::set_error(AMCFD_E_ARG, "cannot start batch threads: " + start_error);
    for (int i = 0; i < n; ++i)
        if (rc[i] < 0) return amcfd::set_error(rc[i], "batch instance " + std::to_string(i) + ": " + msg[i]);
    return AMCFD_OK;
}

int amcfd_run_grouped(amcfd_solver *const *s, int n, int max_steps, int *steps_done) {
    if (n < 0 || (n > 0 && (!s || !steps_done))) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    std::vector<amcfd::Solver *> v(n);
    for (int i = 0; i < n; ++i) {
        if (!s[i]) return amcfd::set_error(AMCFD_E_ARG, "null solver in group");
        v[i] = &s[i]->s;
    }
    std::vector<amcfd::Solver *> seen(v);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
        return amcfd::set_error(AMCFD_E_ARG, "the same solver appears twice in the group");
    GUARD(amcfd::run_grouped(v, max_steps, steps_done));
}

int amcfd_get_field(amcfd_solver *s, const char *name, double *host, size_t n) {
    if (!s || !name || !host) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(s->s.get_field(name, host, n));
}

int amcfd_set_field(amcfd_solver *s, const char *name, const double *host, size_t n) {
    if (!s || !name || !host) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(s->s.set_field(name, host, n));
}

int amcfd_get_geometry(amcfd_solver *s, const char *name, double *host, size_t n) {
    if (!s || !name || !host) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(s->s.get_geometry(name, host, n));
}

int amcfd_get_scalars(amcfd_solver *s, amcfd_scalars *out) {
    if (!s || !out) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(s->s.get_scalars(out));
}

int amcfd_dims(amcfd_solver *s, int *ni, int *nj, int *nk) {
    if (!s || !ni || !nj || !nk) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    s->s.dims(ni, nj, nk);
    return AMCFD_OK;
}

int amcfd_last_timing(amcfd_solver *s, double *total_ms, double *tdma_ms) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (total_ms) *total_ms = s->s.last_total_ms;
    if (tdma_ms) *tdma_ms = s->s.last_tdma_ms;
    return AMCFD_OK;
}

int amcfd_set_profiling(amcfd_solver *s, int on) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (on < 0 || on > 2) return amcfd::set_error(AMCFD_E_ARG, "profiling level must be 0, 1 or 2");
    s->s.set_profiling(on);
    return AMCFD_OK;
}

int amcfd_kernel_stats(amcfd_solver *s, const char *kernel, long *launches, double *total_ms) {
    if (!s || !kernel) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    auto it = s->s.timers.find(kernel);
    if (launches) *launches = it == s->s.timers.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == s->s.timers.end() ? 0.0 : it->second.total_ms;
    return AMCFD_OK;
}

int amcfd_write_vts(amcfd_solver *s, const char *path, const char *const *scalar_names, int n_names) {
    if (!s || !path || (n_names > 0 && !scalar_names)) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(amcfd::write_vts(s->s, path, scalar_names, n_names));
}

int amcfd_write_pvd(const char *path, int n, const double *times, const char *const *files) {
    if (!path || n < 0 || (n > 0 && (!times || !files))) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(amcfd::write_pvd(path, n, times, files));
}

int amcfd_append_residuals(amcfd_solver *s, const char *path) {
    if (!s || !path) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(amcfd::append_residuals(s->s, path));
}

int amcfd_append_final_geometry(amcfd_solver *s, const char *path) {
    if (!s || !path) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(amcfd::append_final_geometry(s->s, path));
}

int amcfd_free_surface(amcfd_solver *s, const char *monitor_path) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(s->s.free_surface(monitor_path));
}

int amcfd_output_section(amcfd_solver *s, const char *base, int finish, int *written) {
    if (!s || !base || !written) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    GUARD(s->s.output_section(base, finish, written));
}

const char *amcfd_deck_string(const amcfd_deck *deck, const char *key, int index) {
    if (!deck || !key) return nullptr;
    const Deck &d = deck->d;
    const std::string k = key;
    const std::string *v = nullptr;
    if (k == "keyword_id") v = &d.keyword_id;
    else if (k == "directory") v = &d.directory;
    else if (k == "final_geometry") v = &d.geometry_out;
    else if (k == "mesh_file") v = &d.mesh_file;
    else if (k == "scalar_out" || k == "vector_out") {
        const std::vector<std::string> &list = k == "scalar_out" ? d.scalar_out : d.vector_out;
        if (index < 0 || index >= (int)list.size()) return nullptr;
        return list[index].c_str();
    }
    if (!v || v->empty()) return nullptr;
    return v->c_str();
}

int amcfd_set_mixed_schedule(amcfd_solver *s, const char *schedule, int box_passes) {
    if (!s || !schedule) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    if (s->s.mode() != AMCFD_MODE_MIXED) return amcfd::set_error(AMCFD_E_ARG, "the solver is not in the mixed mode");
    GUARD(s->s.set_mixed_schedule(schedule, box_passes));
}

int amcfd_set_graph(amcfd_solver *s, int on, int block, int box_blocks) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (s->s.mode() != AMCFD_MODE_MIXED) return amcfd::set_error(AMCFD_E_ARG, "the solver is not in the mixed mode");
    GUARD(s->s.set_graph(on, block, box_blocks));
}

int amcfd_get_graph(const amcfd_solver *s, int *on, int *block, int *box_blocks, int *box_fused, int *line_batch,
                    int *overlap) {
    if (!s || !on || !block || !box_blocks || !box_fused || !line_batch || !overlap)
        return amcfd::set_error(AMCFD_E_ARG, "null argument");
    if (s->s.mode() != AMCFD_MODE_MIXED) return amcfd::set_error(AMCFD_E_ARG, "the solver is not in the mixed mode");
    s->s.graph_settings(on, block, box_blocks);
    *box_fused = s->s.box_fused();
    *line_batch = s->s.line_batch();
    *overlap = s->s.assembly_overlap();
    return AMCFD_OK;
}

int amcfd_set_line_batch(amcfd_solver *s, int lines) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (s->s.mode() != AMCFD_MODE_MIXED) return amcfd::set_error(AMCFD_E_ARG, "the solver is not in the mixed mode");
    GUARD(s->s.set_line_batch(lines));
}

int amcfd_set_assembly_overlap(amcfd_solver *s, int on) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (s->s.mode() != AMCFD_MODE_MIXED) return amcfd::set_error(AMCFD_E_ARG, "the solver is not in the mixed mode");
    GUARD(s->s.set_assembly_overlap(on));
}

int amcfd_set_box_fused(amcfd_solver *s, int on) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (s->s.mode() != AMCFD_MODE_MIXED) return amcfd::set_error(AMCFD_E_ARG, "the solver is not in the mixed mode");
    GUARD(s->s.set_box_fused(on));
}

int amcfd_copy_state(amcfd_solver *dst, const amcfd_solver *src) {
    if (!dst || !src) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    if (dst == src) return AMCFD_OK;
    GUARD(dst->s.copy_state_from(src->s));
}

int amcfd_test_expire_waits(amcfd_solver *s) {
    if (!s) return amcfd::set_error(AMCFD_E_ARG, "null solver");
    GUARD(s->s.raise_sweep_error());
}

int amcfd_test_libm(int fn, const double *x, double *y, size_t n) {
    if ((fn != 0 && fn != 1) || (n && (!x || !y))) return amcfd::set_error(AMCFD_E_ARG, "fn must be 0 or 1, buffers non-null");
    if (amcfd::eval_libm_device(fn, x, y, n)) return amcfd::set_error(AMCFD_E_HIP, "device evaluation failed");
    return AMCFD_OK;
}

int amcfd_ws_attach(amcfd_solver *s, amcfd_ws *ws, int mode) {
    if (!s || (mode && !ws)) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    if (mode < 0 || mode > 3) return amcfd::set_error(AMCFD_E_ARG, "mode must be 0, 1, 2 or 3");
    s->s.attach_warm_start(amcfd::ws_of(ws), mode);
    return AMCFD_OK;
}

int amcfd_ws_counts(const amcfd_solver *s, int *applied, int *guarded) {
    if (!s || !applied || !guarded) return amcfd::set_error(AMCFD_E_ARG, "null argument");
    s->s.warm_start_counts(applied, guarded);
    return AMCFD_OK;
}

const char *amcfd_last_error(void) { return amcfd::g_last_error.c_str(); }

const char *amcfd_version(void) { return "amcfd-mi355x 0.2 (gfx950; FP64 reference-order, scan and mixed modes)"; }

}  // extern "C"
