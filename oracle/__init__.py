"""ctypes access to the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and by the
benchmark's CPU-baseline leg as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

_lib = None


def build(quiet=True):
    """Compile liboracle.so (and the reference harness when the reference
    checkout is present)."""
    out = None if not quiet else subprocess.DEVNULL
    subprocess.run(["make", "-C", _HERE, "liboracle.so"], check=True, stdout=out)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-C", _HERE, "ref"], check=True, stdout=out)


def load(path=LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [vp]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_begin_step.argtypes = [vp]
    lib.orc_solve.argtypes = [vp]
    lib.orc_solve.restype = ctypes.c_int
    lib.orc_end_step.argtypes = [vp]
    lib.orc_run.argtypes = [vp, ctypes.c_int]
    lib.orc_run.restype = ctypes.c_int
    lib.orc_finished.argtypes = [vp]
    lib.orc_finished.restype = ctypes.c_int
    lib.orc_field.argtypes = [vp, ctypes.c_char_p]
    lib.orc_field.restype = dp
    lib.orc_get_scalars.argtypes = [vp, vp]
    lib.orc_free_surface.argtypes = [vp]
    lib.orc_free_surface.restype = ctypes.c_int
    lib.orc_free_surface_monitor.argtypes = [vp] + [ctypes.POINTER(ctypes.c_double)] * 5 + [
        ctypes.POINTER(ctypes.c_int)]
    lib.orc_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int)]
    lib.orc_disc_power.restype = ctypes.c_double
    lib.orc_disc_power.argtypes = [ctypes.c_double]
    lib.orc_disc_exponential.restype = ctypes.c_double
    lib.orc_disc_exponential.argtypes = [ctypes.c_double]
    lib.orc_h_to_t.restype = ctypes.c_double
    lib.orc_h_to_t.argtypes = [vp, ctypes.c_double, dp]
    lib.orc_thomas.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp]
    lib.orc_energy_iteration.argtypes = [vp]
    _lib = lib
    return lib


class Oracle:
    """Sequential FP64 restatement of the per-step solve (reference order)."""

    def __init__(self, params):
        lib = load()
        self._params = params            # keep the struct (and its tables) alive
        self._h = lib.orc_create(ctypes.byref(params))
        ni, nj, nk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.orc_dims(self._h, ctypes.byref(ni), ctypes.byref(nj), ctypes.byref(nk))
        self.shape = (nk.value, nj.value, ni.value)

    def update_laser(self):
        load().orc_begin_step(self._h)

    def iterative_loop_solve(self):
        return load().orc_solve(self._h) == 0

    def update_solutions(self):
        load().orc_end_step(self._h)

    def step(self):
        self.update_laser()
        self.iterative_loop_solve()
        self.update_solutions()
        return bool(load().orc_finished(self._h))

    def run(self, max_steps=-1):
        return load().orc_run(self._h, max_steps)

    def finished(self):
        return bool(load().orc_finished(self._h))

    def field(self, name, shape=None):
        p = load().orc_field(self._h, name.encode())
        if not p:
            raise KeyError(name)
        n = int(np.prod(shape if shape is not None else self.shape))
        return np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(shape if shape is not None else self.shape)

    def scalars(self):
        import importlib
        amcfd = importlib.import_module("am-cfd_amd")
        s = amcfd.Scalars()
        load().orc_get_scalars(self._h, ctypes.byref(s))
        return s.as_dict()

    def free_surface(self):
        """FreeSurfaceManager::execute; returns the monitor values
        (lambda, residual, min fi, max fi, droplet volume, bisections)."""
        lib = load()
        lib.orc_free_surface(self._h)
        d = [ctypes.c_double() for _ in range(5)]
        n = ctypes.c_int()
        lib.orc_free_surface_monitor(self._h, *[ctypes.byref(v) for v in d], ctypes.byref(n))
        return tuple(v.value for v in d) + (n.value,)

    def energy_iteration(self):
        load().orc_energy_iteration(self._h)

    def close(self):
        if self._h:
            load().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
