"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
the twoscalerich_amd package.  Builds nothing: `make -C oracle` (or
__graft_entry__.build()) produces liboracle.so beside this file.
"""
import ctypes as C
import os

import numpy as np

from twoscalerich_amd import _ffi as F

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def load():
    lib = C.CDLL(os.path.join(_HERE, "liboracle.so"))
    lib.oracle_prepare.restype = C.c_void_p
    lib.oracle_prepare.argtypes = [C.POINTER(F.tsr_batch_view)]
    lib.oracle_release.argtypes = [C.c_void_p]
    lib.oracle_gfunc.restype = C.c_int
    lib.oracle_gfunc.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
    lib.oracle_water_balance.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    lib.oracle_vg.argtypes = [C.POINTER(F.soil_struct), _dp, C.c_int, C.c_double, _dp, _dp, _dp]
    lib.oracle_keff.argtypes = [C.POINTER(F.Keff_struct), _dp, C.c_int, _dp]
    lib.oracle_register.argtypes = [C.POINTER(F.user_data_struct), C.c_void_p]
    lib.Gfunc_serial.restype = C.c_int
    lib.Gfunc_serial.argtypes = [_dp, _dp, F.user_data_struct]
    lib.Gfunc_parallel.restype = C.c_int
    lib.Gfunc_parallel.argtypes = [_dp, _dp, F.user_data_struct]
    lib.oracle_phipade.restype = C.c_int
    lib.oracle_phipade.argtypes = [_dp, _dp, C.c_double, C.c_int, C.c_int]
    lib.oracle_expm.restype = C.c_int
    lib.oracle_expm.argtypes = [_dp, _dp, C.c_int]
    lib.oracle_phik_e1.restype = C.c_int
    lib.oracle_phik_e1.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int]
    lib.oracle_weighted_norm.restype = C.c_double
    lib.oracle_weighted_norm.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_longlong]
    lib.oracle_exprem.restype = C.c_int
    lib.oracle_exprem.argtypes = [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp, _dp, F.user_data_struct,
                                  F.exprem_options_struct, C.POINTER(F.exprem_stats_struct)]
    lib.oracle_exprem_last.argtypes = [_dp]
    return lib


class Oracle:
    """CPU FP64 restatement of G(u) and EXPREM over a batch view."""

    def __init__(self, view_ptr, ud_ptr=None):
        self.lib = load()
        self.view = view_ptr
        self.h = self.lib.oracle_prepare(view_ptr)
        self.N = view_ptr.contents.N
        if ud_ptr is not None:
            self.lib.oracle_register(ud_ptr, self.h)

    def close(self):
        if self.h:
            self.lib.oracle_release(self.h)
            self.h = None

    def gfunc(self, u, nthreads=0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        g = np.empty_like(u)
        nt = nthreads or (os.cpu_count() or 1)
        rc = self.lib.oracle_gfunc(self.h, _ptr(g), _ptr(u), nt)
        return g, rc

    def water_balance(self, u, g=None):
        u = np.ascontiguousarray(u, dtype=np.float64)
        w, b, sg = C.c_double(), C.c_double(), C.c_double()
        gp = _ptr(np.ascontiguousarray(g, dtype=np.float64)) if g is not None else None
        self.lib.oracle_water_balance(self.h, _ptr(u), gp, C.byref(w), C.byref(b), C.byref(sg))
        return w.value, b.value, sg.value

    def exprem_step(self, ud, u, t, tend, opt, stats, parallel=True, tspan=None, tspan_solns=None):
        tt = C.c_double(t)
        fn = self.lib.Gfunc_parallel if parallel else self.lib.Gfunc_serial
        fp = C.cast(fn, C.c_void_p)
        rc = self.lib.oracle_exprem(fp, int(self.N), C.byref(tt), float(tend),
                                    _ptr(tspan) if tspan is not None else None,
                                    _ptr(tspan_solns) if tspan_solns is not None else None,
                                    _ptr(u), ud, opt, C.byref(stats))
        last = np.zeros(8)
        self.lib.oracle_exprem_last(_ptr(last))
        return rc, tt.value, last


def phipade(H, alpha, m, maxk):
    lib = load()
    Hc = np.asfortranarray(H, dtype=np.float64)
    assert Hc.shape == (maxk + 1, maxk)
    out = np.zeros((maxk, maxk), order="F")
    rc = lib.oracle_phipade(out.ctypes.data_as(_dp), Hc.ctypes.data_as(_dp), alpha, m, maxk)
    return out[:m, :m], rc


def phik_e1(H, tau, kmax):
    lib = load()
    Hc = np.asfortranarray(H, dtype=np.float64)
    m = Hc.shape[1]
    out = np.zeros(kmax * m)
    rc = lib.oracle_phik_e1(_ptr(out), Hc.ctypes.data_as(_dp), Hc.shape[0], m, tau, kmax)
    return out.reshape(kmax, m), rc


def expm(A):
    lib = load()
    Ac = np.asfortranarray(A, dtype=np.float64)
    E = np.zeros_like(Ac, order="F")
    rc = lib.oracle_expm(E.ctypes.data_as(_dp), Ac.ctypes.data_as(_dp), Ac.shape[0])
    return E, rc
