"""The C-ABI library loads on a machine without a GPU and exports every function that
include/tsr/tsr.h declares (no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import twoscalerich_amd as T
from twoscalerich_amd import _ffi as F

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "tsr", "tsr.h")


def _declared_c_functions():
    src = open(HDR).read()
    c_part = src.split("#ifdef __cplusplus\n}")[0]
    c_part = re.sub(r"/\*.*?\*/", "", c_part, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", c_part)
    return sorted({n for n in names if n not in ("typedef",) and not n.startswith("(")})


def test_every_declared_c_function_is_exported():
    lib = T.lib()
    names = _declared_c_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_cpp_linkage_reference_signatures_are_exported():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", F.LIB_PATH], capture_output=True, text=True).stdout
    for sig in ("macro_mesh_properties(char const*, macro_mesh_struct&, int*)",
                "micro_mesh_properties(char const*, micro_mesh_struct&)",
                "read_effective_conductivity(char const*, Keff_struct&)",
                "set_exprem_options(exprem_options_struct&)",
                "exprem(int (*)(double*, double*, user_data_struct), int, double&, double, double*, double*, double*, "
                "user_data_struct, exprem_options_struct, exprem_stats_struct&)"):
        assert sig in out, sig


def test_ctypes_declarations_cover_the_library():
    assert not F.declare(T.lib())


def test_struct_sizes_match_c_layout():
    # the reference's structs on x86-64 (LP64)
    assert C.sizeof(F.macro_mesh_struct) == 3 * 4 + 4 + 13 * 8
    assert C.sizeof(F.soil_struct) == 48
    assert C.sizeof(F.exprem_options_struct) == 8 * 2 + 4 * 2 + 8 * 2 + 8 + 8 * 4
    assert C.sizeof(F.exprem_stats_struct) == 8 + 8 + 8 + 8 + 4 + 4


def test_host_dense_routines_and_helpers():
    L = T.lib()
    assert L.check_flag_Gfunc(0, 0.5) == 0.5 and L.check_flag_Gfunc(3, 0.5) == 0.25
    o = F.exprem_options_struct()
    L.set_exprem_options_c(C.byref(o))
    assert (o.AbsTol, o.RelTol, o.MinKrylov, o.MaxKrylov, o.MinStep, o.MaxStep, o.MaxNumSteps, o.SafetyFac,
            o.MinStepDecFac, o.MaxStepIncFac, o.InitStep) == (1e-5, 1e-5, 2, 100, 1e-10, 1000.0, 100000, 0.8, 0.1,
                                                            1.1, 1e-3)


def test_solver_create_rejects_bad_options_without_a_gpu():
    # checked before any device work: NULL and a message; the partly built solver is
    # released
    import pytest
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    o = p.options()
    o.MaxKrylov = 0
    assert T.lib().tsr_solver_create(p.view, C.byref(o)) is None
    assert b"MaxKrylov" in T.lib().tsr_last_error()
    with pytest.raises(RuntimeError, match="MaxKrylov"):
        T.Solver(p, o)
    p.close()


def test_gfunc_hip_rejects_unregistered_user_data_of_unknown_shape():
    # a user_data_struct no set-up call registered is taken as the reference's own problem
    # only when N fits one of its two hosting rules; otherwise Gfunc_hip returns 2 with the
    # reason, before any device work
    import numpy as np
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    ud = F.user_data_struct.from_buffer_copy(p.ud.contents)
    ud.N += 1
    u = np.zeros(ud.N)
    g = np.zeros(ud.N)
    assert T.lib().Gfunc_hip(T._dp(g), T._dp(u), ud) == 2
    assert b"M_nodes + M_elements*m_nodes" in T.lib().tsr_last_error()
    p.close()


def test_exprem_c_rejects_bad_options_without_a_gpu():
    # the solver exprem_c makes for a problem is checked before any device work; a
    # failed set-up returns -1 (and does not hold the solver cache's lock)
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    o = p.options()
    o.MaxKrylov = 500
    u = p.initial()
    t = C.c_double(0.0)
    st = F.exprem_stats_struct()
    gf = C.cast(T.lib().Gfunc_hip, C.c_void_p)
    for _ in range(2):
        rc = T.lib().exprem_c(gf, p.N, C.byref(t), p.tend(), None, None, T._dp(u), p.ud.contents, o, C.byref(st))
        assert rc == -1
    assert b"MaxKrylov" in T.lib().tsr_last_error()
    p.close()


def test_unregistered_user_data_hosting_rules_without_a_gpu():
    # N = M_nodes * (1 + m_nodes) is taken as cells hosted by macro nodes (this workload's
    # layout) and N = M_nodes + M_elements * m_nodes as cells hosted by macro elements: both
    # pass the shape check and reach the device upload, which is what fails without a GPU
    import numpy as np
    import torch
    if torch.cuda.device_count() > 0:
        import pytest
        pytest.skip("checks the failure past the shape check, which needs a machine without a GPU")
    p = T.Problem.config(F.CFG_PLAIN_SMALL)
    ud = F.user_data_struct.from_buffer_copy(p.ud.contents)
    M, m = ud.macro_mesh, ud.micro_mesh
    el = np.ctypeslib.as_array(M.elements, shape=(M.no_verts * M.no_elements,)).copy()
    ud.macro_mesh.elements = el.ctypes.data_as(C.POINTER(C.c_int))
    for N in (M.no_nodes * (1 + m.no_nodes), M.no_nodes + M.no_elements * m.no_nodes):
        ud.N = N
        u = np.zeros(N)
        g = np.zeros(N)
        assert T.lib().Gfunc_hip(T._dp(g), T._dp(u), ud) == 2
        assert b"neither" not in T.lib().tsr_last_error()
    p.close()
