"""MI355X-native TwoScaleRich time-stepping hot path.

Host-side handles over libtsr.so (HIP kernels for gfx950 + C++ host code):

    Problem   a workload (a benchmark case or the reference's input files) as a
              batch view, with the reference's user_data_struct for drop-in calls;
    Solver    the device-resident EXPREM integrator and the GPU right-hand side.

The library is required: importing works without it (so the CPU test suite can run),
but every compute call raises if libtsr.so is missing or failed to load.  There is no
CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _ffi as F

_lib = None
_missing = None


def lib():
    """The loaded libtsr.so (raises if it is not built)."""
    global _lib, _missing
    if _lib is None:
        if not os.path.exists(F.LIB_PATH):
            raise RuntimeError(f"{F.LIB_PATH} is not built: run __graft_entry__.build() or "
                               "python -m twoscalerich_amd.build")
        L = C.CDLL(F.LIB_PATH)
        _missing = F.declare(L)
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def last_error():
    return (lib().tsr_last_error() or b"").decode()


class Problem:
    """A workload assembled by the library's set-up code (no compute)."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("problem set-up failed: " + last_error())
        self.h = handle
        L = lib()
        self.view = L.tsr_problem_view(handle)
        self.ud = L.tsr_problem_user_data(handle)
        self.N = int(L.tsr_problem_N(handle))

    @classmethod
    def config(cls, cfg, macro_n=0, micro_r=0, n_scen=0, seed=0):
        return cls(lib().tsr_problem_config(cfg, macro_n, micro_r, n_scen, seed))

    @classmethod
    def reference(cls, macro_msh, micro_msh, keff_txt, soilA, soilB, boundary_flux, macro_bc, model_type, h0,
                  host_kind=F.HOST_ELEMENT):
        bc = (C.c_int * 5)(*macro_bc)
        h = lib().tsr_problem_reference(macro_msh.encode(), micro_msh.encode(), keff_txt.encode(), C.byref(soilA),
                                        C.byref(soilB), boundary_flux, bc, model_type, h0, host_kind)
        return cls(h)

    def initial(self):
        u = np.empty(self.N)
        lib().tsr_problem_initial(self.h, _dp(u))
        return u

    def tend(self):
        return lib().tsr_problem_tend(self.h)

    def options(self):
        o = F.exprem_options_struct()
        lib().tsr_problem_options(self.h, C.byref(o))
        return o

    def distinct_bytes(self):
        return int(lib().tsr_problem_distinct_bytes(self.h))

    def close(self):
        if self.h:
            lib().tsr_problem_free(self.h)
            self.h = None


class Solver:
    """Device-resident EXPREM over a Problem's batch view."""

    def __init__(self, problem, options=None):
        self.problem = problem
        self.opt = options if options is not None else problem.options()
        self.h = lib().tsr_solver_create(problem.view, C.byref(self.opt))
        if not self.h:
            raise RuntimeError("tsr_solver_create failed: " + last_error())
        self.N = problem.N

    def gfunc(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        g = np.empty_like(u)
        rc = lib().tsr_solver_gfunc_host(self.h, _dp(g), _dp(u))
        if rc == 2:
            raise RuntimeError("Gfunc on the device failed: " + last_error())
        return g, rc

    def set_state(self, u, t=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if lib().tsr_solver_set_state(self.h, _dp(u), t):
            raise RuntimeError(last_error())

    def get_state(self):
        u = np.empty(self.N)
        t = C.c_double()
        if lib().tsr_solver_get_state(self.h, _dp(u), C.byref(t)):
            raise RuntimeError(last_error())
        return u, t.value

    def reset(self):
        lib().tsr_solver_reset(self.h)

    def step(self, tend):
        S = self.problem.view.contents.n_scen
        rec = (F.tsr_step_record * S)()
        rc = lib().tsr_solver_step(self.h, tend, rec)
        return rc, list(rec)

    def run(self, tend, max_steps=10 ** 7, max_records=0):
        st = F.tsr_run_stats()
        recs = (F.tsr_step_record * max_records)() if max_records else None
        rc = lib().tsr_solver_run(self.h, tend, max_steps, C.byref(st), recs, max_records)
        return rc, st, (list(recs) if recs is not None else [])

    def stats(self):
        s = F.exprem_stats_struct()
        lib().tsr_solver_stats(self.h, C.byref(s))
        return s

    def set_path(self, path):
        if lib().tsr_solver_set_path(self.h, path):
            raise RuntimeError("tsr_solver_set_path failed: " + last_error())

    def jv(self, v, eps, lin=False):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.empty_like(v)
        if lib().tsr_solver_jv(self.h, _dp(v), eps, 1 if lin else 0, _dp(out)):
            raise RuntimeError("tsr_solver_jv failed: " + last_error())
        return out

    def set_controller(self, mode):
        return lib().tsr_solver_set_controller(self.h, mode)

    def set_step(self, tau):
        if lib().tsr_solver_set_step(self.h, tau):
            raise ValueError(f"step size {tau!r} must be positive and finite")

    def ktimer(self, every):
        """Sample every `every`-th launch of each timed kernel class with HIP events (0: off)."""
        if lib().tsr_solver_ktimer(self.h, every):
            raise ValueError("every must be >= 0")

    def ktimer_read(self):
        """Per kernel class (_ffi.KT_NAMES): launches seen and timed, device ms and
        algorithmic bytes of the timed launches."""
        n = F.KT_COUNT
        seen, timed = (C.c_longlong * n)(), (C.c_longlong * n)()
        ms, by = (C.c_double * n)(), (C.c_double * n)()
        lib().tsr_solver_ktimer_read(self.h, seen, timed, ms, by)
        return {F.KT_NAMES[c]: {"seen": seen[c], "timed": timed[c], "ms": ms[c], "bytes": by[c]} for c in range(n)}

    def close(self):
        if self.h:
            lib().tsr_solver_free(self.h)
            self.h = None
