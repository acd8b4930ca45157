"""Python mirror of AM-CFD's per-time-step solver surface over ``libamcfd.so``.

The reference drives its solve through C++ member functions
(src/main.cpp:56-162):

    Mesh::getDomainInfo / readMeshFile        -> Deck(path)
    BoundCondManager::updateLaser             -> Solver.update_laser()
    CFDSolverManager::iterativeLoopSolve      -> Solver.iterative_loop_solve()
    CFDSolverManager::updateSolutions         -> Solver.update_solutions()
    SolutionVariableManager::mapSolutionVars_ -> Solver.field(name)

Everything below is a thin ctypes layer over the C ABI in include/amcfd.h;
all arithmetic runs in the HIP kernels of libamcfd.so.  There is no CPU
fallback: constructing a Solver without the built library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamcfd.so")
CLI_PATH = os.path.join(_HERE, "amcfd")

MAX_ZONES = 32
OK, I_MAXIT = 0, 1
MODE_FP64 = 0
MODE_SCAN = 1
MODE_MIXED = 2

FIELD_NAMES = ("enthalpy", "temp", "fracl", "uvel", "vvel", "wvel", "pressure", "pp", "rho",
               "diff", "vis", "hnot", "fraclnot", "unot", "vnot", "wnot", "sourceinput", "dux",
               "dvy", "dwz", "csfrac", "solfrac", "mpPhase", "mpLiqPhase")

# header symbols of include/amcfd.h (checked by the CPU tests)
EXPORTS = ("amcfd_deck_load", "amcfd_deck_params", "amcfd_deck_free", "amcfd_create",
           "amcfd_destroy", "amcfd_begin_step", "amcfd_solve_step", "amcfd_end_step",
           "amcfd_step", "amcfd_get_field", "amcfd_set_field", "amcfd_get_geometry",
           "amcfd_get_scalars", "amcfd_dims", "amcfd_last_timing", "amcfd_set_profiling",
           "amcfd_kernel_stats", "amcfd_write_vts", "amcfd_write_pvd", "amcfd_append_residuals",
           "amcfd_append_final_geometry", "amcfd_deck_string", "amcfd_free_surface",
           "amcfd_output_section", "amcfd_run_batch", "amcfd_run_grouped", "amcfd_set_mixed_schedule", "amcfd_set_graph", "amcfd_get_graph", "amcfd_set_box_fused", "amcfd_set_line_batch", "amcfd_set_assembly_overlap",
           "amcfd_tuning_choice", "amcfd_copy_state", "amcfd_test_expire_waits", "amcfd_test_libm",
           "amcfd_tdma_lines", "amcfd_tdma_lines_noalloc",
           "amcfd_ws_create", "amcfd_ws_destroy", "amcfd_ws_attach", "amcfd_ws_counts", "amcfd_ws_train", "amcfd_ws_samples",
           "amcfd_ws_set_factors", "amcfd_ws_save", "amcfd_ws_load", "amcfd_ws_param_count", "amcfd_ws_params",
           "amcfd_ws_loss_grad",
           "amcfd_last_error", "amcfd_version")


class Params(ctypes.Structure):
    """ctypes mirror of ``amcfd_params``."""
    _Z = ctypes.c_double * MAX_ZONES
    _fields_ = [
        ("nzx", ctypes.c_int), ("nzy", ctypes.c_int), ("nzz", ctypes.c_int),
        ("xzone", _Z), ("ncvx", _Z), ("powrx", _Z),
        ("yzone", _Z), ("ncvy", _Z), ("powry", _Z),
        ("zzone", _Z), ("ncvz", _Z), ("powrz", _Z),
    ] + [(n, ctypes.c_double) for n in (
        "finaltime", "maxit", "nonlintol", "delt", "outtime",
        "grav", "sigm", "tempamb", "pressamb", "gasconst")] + [
        (n, ctypes.c_int) for n in (
            "navierstokes", "discretization", "steady", "isevap", "powderbed", "udtoolpath",
            "gaussheatsource", "volheatsource", "volheatsourcetable")] + [
        (n, ctypes.c_double) for n in (
            "lasertime", "xstart", "ystart", "scanvel",
            "dens", "denl", "viscos", "tsolid", "tliquid", "hsmelt", "hlfriz", "acpa", "acpb", "acpl",
            "thconsa", "thconsb", "thconsc", "thconla", "thconlb",
            "urfu", "urfv", "urfw", "urfp", "urfh",
            "htci", "htcj", "htck1", "htckn", "emiss",
            "tempwest", "tempeast", "tempnorth", "tempbottom", "temppreheat",
            "beta", "dgdtp", "darcyKo", "a_oxygen",
            "alaspow", "alaseta", "alasrb", "alasfact",
            "evaplatent", "evaptemp", "evapfact", "molarmass",
            "avolpow", "apowseta", "heatthick", "heatrb", "avolfact")] + [
        ("numlayers", ctypes.c_int), ("ncvzlayer", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in (
            "layerheight", "powderden", "powdercpa", "powdercpb", "powderthcona", "powderthconb")] + [
        ("n_toolpath", ctypes.c_int), ("toolpath", ctypes.POINTER(ctypes.c_double)),
        ("n_dt_table", ctypes.c_int), ("dt_table", ctypes.POINTER(ctypes.c_double)),
        ("n_heat_table", ctypes.c_int), ("heat_table", ctypes.POINTER(ctypes.c_double)),
        ("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nk", ctypes.c_int), ("nk_active0", ctypes.c_int),
        ("hlcal", ctypes.c_double), ("hlatnt", ctypes.c_double),
        ("movingmesh", ctypes.c_int), ("xmoveindex", ctypes.c_int), ("ymoveindex", ctypes.c_int),
        ("species", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("initconcentration", "massdiff", "kp", "urfc")] + [
        ("solidification", ctypes.c_int), ("energyfreesurface", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("initlam1", "initlam2", "surftens", "maxiterenergy", "tolenergy",
                                       "maxiterbisect", "tolbisect", "urfree")] + [
        ("outputsection", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in ("gxmin", "gymin", "gzmin", "gxmax", "gymax", "gzmax", "toutsect")]


class Scalars(ctypes.Structure):
    """ctypes mirror of ``amcfd_scalars``."""
    _fields_ = [("timet", ctypes.c_double), ("niter", ctypes.c_int), ("itertot", ctypes.c_int)] + [
        (n, ctypes.c_double) for n in (
            "resorh", "resorm", "resoru", "resorv", "resorw", "tpeak", "alen", "depth", "width",
            "ratio", "heatinlaser", "heatout", "accul", "heatvol", "ahtoploss", "beamposx",
            "beamposy", "scanvel")] + [
        (n, ctypes.c_int) for n in (
            "powerindicator", "trackindx", "istat", "iend", "jstat", "jend", "kstat", "istart",
            "jstart", "nk", "step")] + [("delt", ctypes.c_double), ("outtime", ctypes.c_double),
                                        ("resorc", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class AmcfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("amcfd error %d: %s" % (code, msg))
        self.code = code


_lib = None


def load_library(path=LIB_PATH):
    """Load libamcfd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libamcfd.so is not built (%s); run __graft_entry__.build()" % path)
    lib = ctypes.CDLL(path)
    vp, ip, dp, cp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p
    sig = {
        "amcfd_deck_load": (ip, [cp, ctypes.POINTER(vp)]),
        "amcfd_deck_params": (ctypes.POINTER(Params), [vp]),
        "amcfd_deck_free": (None, [vp]),
        "amcfd_create": (ip, [ctypes.POINTER(Params), ip, ctypes.POINTER(vp)]),
        "amcfd_destroy": (None, [vp]),
        "amcfd_begin_step": (ip, [vp]),
        "amcfd_solve_step": (ip, [vp]),
        "amcfd_end_step": (ip, [vp]),
        "amcfd_step": (ip, [vp, ctypes.POINTER(ip)]),
        "amcfd_get_field": (ip, [vp, cp, dp, ctypes.c_size_t]),
        "amcfd_set_field": (ip, [vp, cp, dp, ctypes.c_size_t]),
        "amcfd_get_geometry": (ip, [vp, cp, dp, ctypes.c_size_t]),
        "amcfd_get_scalars": (ip, [vp, ctypes.POINTER(Scalars)]),
        "amcfd_dims": (ip, [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), ctypes.POINTER(ip)]),
        "amcfd_last_timing": (ip, [vp, dp, dp]),
        "amcfd_set_profiling": (ip, [vp, ip]),
        "amcfd_kernel_stats": (ip, [vp, cp, ctypes.POINTER(ctypes.c_long), dp]),
        "amcfd_write_vts": (ip, [vp, cp, ctypes.POINTER(cp), ip]),
        "amcfd_write_pvd": (ip, [cp, ip, dp, ctypes.POINTER(cp)]),
        "amcfd_append_residuals": (ip, [vp, cp]),
        "amcfd_append_final_geometry": (ip, [vp, cp]),
        "amcfd_deck_string": (cp, [vp, cp, ip]),
        "amcfd_free_surface": (ip, [vp, cp]),
        "amcfd_output_section": (ip, [vp, cp, ip, ctypes.POINTER(ip)]),
        "amcfd_run_batch": (ip, [ctypes.POINTER(vp), ip, ip, ip, ctypes.POINTER(ip)]),
        "amcfd_run_grouped": (ip, [ctypes.POINTER(vp), ip, ip, ctypes.POINTER(ip)]),
        "amcfd_set_mixed_schedule": (ip, [vp, cp, ip]),
        "amcfd_test_expire_waits": (ip, [vp]),
        "amcfd_test_libm": (ip, [ip, dp, dp, ctypes.c_size_t]),
        "amcfd_set_graph": (ip, [vp, ip, ip, ip]),
        "amcfd_get_graph": (ip, [vp] + [ctypes.POINTER(ctypes.c_int)] * 6),
        "amcfd_set_line_batch": (ip, [vp, ip]),
        "amcfd_set_assembly_overlap": (ip, [vp, ip]),
        "amcfd_set_box_fused": (ip, [vp, ip]),
        "amcfd_tuning_choice": (ip, [cp, ctypes.c_longlong] + [ctypes.POINTER(ctypes.c_int)] * 5),
        "amcfd_copy_state": (ip, [vp, vp]),
        "amcfd_tdma_lines": (ip, [vp, vp, vp, vp, vp, ip, ip, ip, vp]),
        "amcfd_tdma_lines_noalloc": (ip, [vp, vp, vp, vp, vp, ip, ip, ip, vp]),
        "amcfd_ws_create": (ip, [ip, ctypes.POINTER(vp)]),
        "amcfd_ws_destroy": (None, [vp]),
        "amcfd_ws_attach": (ip, [vp, vp, ip]),
        "amcfd_ws_counts": (ip, [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "amcfd_ws_train": (ip, [vp, ip, ip, ctypes.c_double, dp, dp]),
        "amcfd_ws_samples": (ip, [vp]),
        "amcfd_ws_set_factors": (ip, [vp, ctypes.c_double, ctypes.c_double]),
        "amcfd_ws_save": (ip, [vp, cp]),
        "amcfd_ws_load": (ip, [vp, cp]),
        "amcfd_ws_param_count": (ip, [vp]),
        "amcfd_ws_params": (ip, [vp, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t, ip]),
        "amcfd_ws_loss_grad": (ip, [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_float), ip, ip, ip, ip, dp,
                                    ctypes.POINTER(ctypes.c_float)]),
        "amcfd_last_error": (cp, []),
        "amcfd_version": (cp, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(code):
    if code < 0:
        raise AmcfdError(code, load_library().amcfd_last_error().decode())
    return code


class Deck:
    """Parsed keyword deck (reference grammar, Mesh.cpp:39-550)."""

    def __init__(self, path):
        lib = load_library()
        self._h = ctypes.c_void_p()
        _check(lib.amcfd_deck_load(os.fsencode(path), ctypes.byref(self._h)))
        # a copy of the struct; its table pointers still refer to this deck,
        # so the copy keeps the deck alive
        self.params = Params.from_buffer_copy(lib.amcfd_deck_params(self._h).contents)
        self.params._deck = self

    def close(self):
        if self._h:
            load_library().amcfd_deck_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Device-resident solver with the reference's per-step call surface."""

    def __init__(self, deck_or_params, mode=MODE_FP64):
        lib = load_library()
        if isinstance(deck_or_params, Deck):
            self._deck = deck_or_params
            params = deck_or_params.params
        elif isinstance(deck_or_params, (str, os.PathLike)):
            self._deck = Deck(deck_or_params)
            params = self._deck.params
        else:
            self._deck = None
            params = deck_or_params
        self.params = params
        self._h = ctypes.c_void_p()
        _check(lib.amcfd_create(ctypes.byref(params), mode, ctypes.byref(self._h)))
        ni, nj, nk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.amcfd_dims(self._h, ctypes.byref(ni), ctypes.byref(nj), ctypes.byref(nk))
        self.shape = (nk.value, nj.value, ni.value)   # numpy (k, j, i) order

    # -- reference call surface -------------------------------------------------
    def update_laser(self):
        """dt table + timet += dt + BoundCondManager::updateLaser."""
        return _check(load_library().amcfd_begin_step(self._h))

    def iterative_loop_solve(self):
        """CFDSolverManager::iterativeLoopSolve; returns True if it converged."""
        return _check(load_library().amcfd_solve_step(self._h)) == OK

    def update_solutions(self):
        """CFDSolverManager::updateSolutions."""
        return _check(load_library().amcfd_end_step(self._h))

    def step(self):
        """One full time step; returns True once timet >= finaltime."""
        fin = ctypes.c_int(0)
        _check(load_library().amcfd_step(self._h, ctypes.byref(fin)))
        return bool(fin.value)

    def run(self, max_steps=None):
        n = 0
        while max_steps is None or n < max_steps:
            n += 1
            if self.step():
                break
        return n

    def free_surface(self, monitor_path=None):
        """Energy-based free surface at an output time (updates "zr", "tempr")."""
        _check(load_library().amcfd_free_surface(self._h, monitor_path.encode() if monitor_path else None))

    def output_section(self, base, finish=False):
        """Extracted-section cadence step; True if a section was written."""
        w = ctypes.c_int(0)
        _check(load_library().amcfd_output_section(self._h, base.encode(), int(finish), ctypes.byref(w)))
        return bool(w.value)

    # -- data access ------------------------------------------------------------
    def field(self, name):
        out = np.empty(self.shape, dtype=np.float64)
        _check(load_library().amcfd_get_field(
            self._h, name.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
        return out

    def set_field(self, name, values):
        arr = np.ascontiguousarray(values, dtype=np.float64).reshape(self.shape)
        _check(load_library().amcfd_set_field(
            self._h, name.encode(), arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), arr.size))

    def geometry(self, name):
        nk, nj, ni = self.shape
        sizes = {"x": ni, "y": nj, "z": nk, "xu": ni, "yv": nj, "zw": nk, "dxpwinv": ni,
                 "dypsinv": nj, "dzpbinv": nk, "fracx": ni, "fracy": nj, "fracz": nk,
                 "areaij": nj * ni, "areajk": nk * nj, "areaik": nk * ni, "volume": nk * nj * ni}
        out = np.empty(sizes[name], dtype=np.float64)
        _check(load_library().amcfd_get_geometry(
            self._h, name.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size))
        return out

    def scalars(self):
        s = Scalars()
        _check(load_library().amcfd_get_scalars(self._h, ctypes.byref(s)))
        return s.as_dict()

    def last_timing(self):
        t, d = ctypes.c_double(), ctypes.c_double()
        load_library().amcfd_last_timing(self._h, ctypes.byref(t), ctypes.byref(d))
        return t.value, d.value

    def set_profiling(self, on=True):
        """Time the kernel families with HIP events: True/1 every family,
        2 the whole-domain enthalpy sweep alone, False/0 none."""
        _check(load_library().amcfd_set_profiling(self._h, 2 if on == 2 else int(bool(on))))

    def kernel_stats(self, kernel):
        n, ms = ctypes.c_long(), ctypes.c_double()
        _check(load_library().amcfd_kernel_stats(self._h, kernel.encode(), ctypes.byref(n),
                                                 ctypes.byref(ms)))
        return n.value, ms.value

    def set_mixed_schedule(self, schedule, box_passes=4):
        """Mixed mode: order of the enthalpy line passes ('x', 'y', 'z') and
        red-black x passes per melt-pool box variable."""
        _check(load_library().amcfd_set_mixed_schedule(self._h, schedule.encode(), int(box_passes)))

    def set_graph(self, on=True, block=8, box_blocks=256):
        """Mixed mode: captured-graph SIMPLE loop (device-side control) or host loop."""
        _check(load_library().amcfd_set_graph(self._h, int(bool(on)), int(block), int(box_blocks)))

    def warm_start_counts(self):
        """(steps started from the warm start's guess, steps its guard kept
        on the plain start)."""
        a, g = ctypes.c_int(), ctypes.c_int()
        _check(load_library().amcfd_ws_counts(self._h, ctypes.byref(a), ctypes.byref(g)))
        return a.value, g.value

    def graph_settings(self):
        """Mixed mode: (captured loop on, iterations per graph, box workgroups,
        box passes in one persistent launch, x-lines per workgroup, assembly
        overlap on)."""
        on, b, bb, bf, lb, ov = (ctypes.c_int() for _ in range(6))
        _check(load_library().amcfd_get_graph(self._h, ctypes.byref(on), ctypes.byref(b), ctypes.byref(bb),
                                              ctypes.byref(bf), ctypes.byref(lb), ctypes.byref(ov)))
        return bool(on.value), b.value, bb.value, bool(bf.value), lb.value, bool(ov.value)

    def set_line_batch(self, lines=1):
        """Mixed mode: consecutive x-lines per workgroup in the enthalpy's
        passes; the same results."""
        _check(load_library().amcfd_set_line_batch(self._h, int(lines)))

    def set_assembly_overlap(self, on=True):
        """Mixed mode, captured loop: the next iteration's assembly away from
        the melt-pool box beside the flow part (True, or a workgroup count
        for it) or after it (False); the same results."""
        _check(load_library().amcfd_set_assembly_overlap(self._h, int(on)))

    def set_box_fused(self, on=True):
        """Mixed mode: a box variable's passes in one persistent launch (True)
        or one launch per half pass (False); the same results."""
        _check(load_library().amcfd_set_box_fused(self._h, int(bool(on))))

    def copy_state_from(self, other):
        """Continue from another solver's exact state (same deck)."""
        _check(load_library().amcfd_copy_state(self._h, other._h))

    def test_expire_waits(self):
        """Raise the line sweeps' device error flag (test hook)."""
        _check(load_library().amcfd_test_expire_waits(self._h))

    def close(self):
        if self._h:
            load_library().amcfd_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batch(solvers, max_steps=-1, workers=0):
    """Advance independent solvers concurrently on the one GPU (worker host
    threads, one stream per instance); returns the steps each took."""
    n = len(solvers)
    handles = (ctypes.c_void_p * max(n, 1))(*[s._h.value for s in solvers])
    done = (ctypes.c_int * max(n, 1))()
    _check(load_library().amcfd_run_batch(handles, n, int(max_steps), int(workers), done))
    return [done[i] for i in range(n)]


def run_grouped(solvers, max_steps=-1):
    """Advance mixed-mode solvers in lock step, each kernel of an iteration
    one launch over all of them (captured); returns the steps each took."""
    n = len(solvers)
    handles = (ctypes.c_void_p * max(n, 1))(*[s._h.value for s in solvers])
    done = (ctypes.c_int * max(n, 1))()
    _check(load_library().amcfd_run_grouped(handles, n, int(max_steps), done))
    return [done[i] for i in range(n)]


def tuning_choice(cells, path=None):
    """(graph_block, box_blocks, box_fused, line_batch, overlap) the mixed
    mode takes for a grid of `cells` interior cells, from the table at `path`
    or the library's own."""
    gb, bb, bf, lb, ov = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(load_library().amcfd_tuning_choice(os.fsencode(path) if path else None, int(cells), ctypes.byref(gb),
                                              ctypes.byref(bb), ctypes.byref(bf), ctypes.byref(lb), ctypes.byref(ov)))
    return gb.value, bb.value, bool(bf.value), lb.value, bool(ov.value)


def device_libm(fn, x):
    """The device's exp (fn "exp") or log (fn "log") of the float64 array x,
    as the kernels evaluate them."""
    import numpy as np
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    code = {"exp": 0, "log": 1}[fn]
    _check(load_library().amcfd_test_libm(code, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          y.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), x.size))
    return y


def tdma_lines(aw, ap, ae, d, phi, n, nlines, guard=0, stream=None):
    """The reference's x-line TDMA on device buffers (device addresses as
    ints); phi is solved in place.  Returns once the work is queued."""
    _check(load_library().amcfd_tdma_lines(aw, ap, ae, d, phi, int(n), int(nlines), int(guard), stream))


def tdma_lines_noalloc(aw, ap, ae, d, phi, n, nlines, guard=0, stream=None):
    """amcfd_tdma_lines with the library's fixed device scratch: allocates
    nothing; n * nlines <= 2**20."""
    _check(load_library().amcfd_tdma_lines_noalloc(aw, ap, ae, d, phi, int(n), int(nlines), int(guard), stream))


# warm start modes: off, record samples, start a step from the guess when the
# step before converged, start every step from the guess
WS_OFF, WS_RECORD, WS_APPLY, WS_APPLY_ALWAYS = 0, 1, 2, 3


class WarmStartNet:
    """The learned warm start (3-D conv encoder-decoder on BF16 matrix cores)."""

    def __init__(self, seed=0):
        self._h = ctypes.c_void_p()
        _check(load_library().amcfd_ws_create(int(seed), ctypes.byref(self._h)))

    def attach(self, solver, mode):
        _check(load_library().amcfd_ws_attach(solver._h, self._h if mode else None, int(mode)))

    def train(self, iters, batch=8, lr=1e-3):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(load_library().amcfd_ws_train(self._h, int(iters), int(batch), float(lr), ctypes.byref(a),
                                             ctypes.byref(b)))
        return a.value, b.value

    def samples(self):
        return _check(load_library().amcfd_ws_samples(self._h))

    def set_factors(self, tfactor=1.0, pfactor=0.0):
        _check(load_library().amcfd_ws_set_factors(self._h, float(tfactor), float(pfactor)))

    def save(self, path):
        _check(load_library().amcfd_ws_save(self._h, os.fsencode(path)))

    def load(self, path):
        _check(load_library().amcfd_ws_load(self._h, os.fsencode(path)))

    def params(self):
        n = _check(load_library().amcfd_ws_param_count(self._h))
        out = np.empty(n, dtype=np.float32)
        _check(load_library().amcfd_ws_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, 0))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(load_library().amcfd_ws_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size, 1))

    def loss_grad(self, x, target, mask):
        """x: (N, D, H, W, 16), target: (N, D, H, W, 2), mask: (N, D, H, W)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        n, d, h, w = m.shape
        loss = ctypes.c_double()
        g = np.empty(_check(load_library().amcfd_ws_param_count(self._h)), dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        _check(load_library().amcfd_ws_loss_grad(self._h, x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                                 m.ctypes.data_as(fp), n, d, h, w, ctypes.byref(loss),
                                                 g.ctypes.data_as(fp)))
        return loss.value, g

    def close(self):
        if self._h:
            load_library().amcfd_ws_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def version():
    return load_library().amcfd_version().decode()
