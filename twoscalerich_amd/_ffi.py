"""ctypes mirror of include/tsr/tsr_types.h and the C-ABI of libtsr.so.

The struct layouts are the reference's plain-C carriers (include/macro_mesh.h,
micro_mesh.h, Keff.h, soil.h, user_data.h, exprem_options.h, exprem_stats.h) plus the
batch view; field order and types must match the C header exactly.
"""
import ctypes as C
import os

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


class macro_mesh_struct(C.Structure):
    _fields_ = [("no_verts", C.c_int), ("no_nodes", C.c_int), ("no_elements", C.c_int),
                ("elements", _ip), ("nodes", _dp), ("shape_funcs", _dp), ("normals", _dp),
                ("boundary_edge_midpts", _dp), ("boundary_edge_normals", _dp), ("boundary_edge_lens", _dp),
                ("boundary_condition_edges", _ip), ("CV_area", _dp), ("SCV_area", _dp), ("midx", _dp),
                ("midy", _dp), ("centroid_elements", _dp)]


class micro_mesh_struct(C.Structure):
    _fields_ = [("no_nodes", C.c_int), ("no_elements", C.c_int), ("elements", _ip), ("boundary_nodes", _ip),
                ("nodes", _dp), ("shape_funcs", _dp), ("normals", _dp), ("boundary_edges", _ip),
                ("boundary_edge_midpts", _dp), ("boundary_edge_normals", _dp), ("boundary_edge_lens", _dp),
                ("centroid_elements", _dp), ("CV_area", _dp), ("midx", _dp), ("midy", _dp),
                ("cell_area", C.c_double), ("vectors", _dp), ("no_verts", C.c_int), ("element_area", _dp)]


class Keff_struct(C.Structure):
    _fields_ = [("epsilonA", C.c_double), ("no_hvalues", C.c_int), ("logr", C.c_double),
                ("logh_first", C.c_double), ("logh_values", _dp), ("Keff_values", _dp)]


class soil_struct(C.Structure):
    _fields_ = [("Ksat", C.c_double), ("thetar", C.c_double), ("thetas", C.c_double),
                ("alpha", C.c_double), ("n", C.c_double), ("m", C.c_double)]


class user_data_struct(C.Structure):
    _fields_ = [("macro_mesh", macro_mesh_struct), ("micro_mesh", micro_mesh_struct),
                ("effective_conductivity", Keff_struct), ("soilA", soil_struct), ("soilB", soil_struct),
                ("boundary_flux", C.c_double), ("N", C.c_int), ("model_type", C.c_int),
                ("NUM_THREADS", C.c_uint)]


class exprem_options_struct(C.Structure):
    _fields_ = [("AbsTol", C.c_double), ("RelTol", C.c_double), ("MinKrylov", C.c_int), ("MaxKrylov", C.c_int),
                ("MinStep", C.c_double), ("MaxStep", C.c_double), ("MaxNumSteps", C.c_int),
                ("SafetyFac", C.c_double), ("MinStepDecFac", C.c_double), ("MaxStepIncFac", C.c_double),
                ("InitStep", C.c_double)]


class exprem_stats_struct(C.Structure):
    _fields_ = [("CurrentStep", C.c_double), ("CurrentTime", C.c_double), ("CurrentKryDim", C.c_int),
                ("CurrentSoln", _dp), ("NumSteps", C.c_int), ("FuncEvals", C.c_int)]


class tsr_scenario(C.Structure):
    _fields_ = [("ud", user_data_struct), ("host_kind", C.c_int), ("n_cells", C.c_int),
                ("first_cell", C.c_int), ("pad_", C.c_int), ("u_offset", C.c_longlong),
                ("dirichlet_head", C.c_double)]


class tsr_batch_view(C.Structure):
    _fields_ = [("n_scen", C.c_int), ("n_tmpl", C.c_int), ("n_cells", C.c_int), ("max_discs", C.c_int),
                ("scen", C.POINTER(tsr_scenario)), ("tmpl", C.POINTER(micro_mesh_struct)),
                ("tmpl_soil", C.POINTER(C.POINTER(C.c_ubyte))), ("cell_tmpl", _ip),
                ("cell_offset", C.POINTER(C.c_longlong)), ("cell_discs", _dp), ("Ss", C.c_double),
                ("N", C.c_longlong)]


class tsr_step_record(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t", "tau", "tau_proposed", "err", "beta", "hnorm", "resid",
                                          "head_mean", "head_min", "head_max", "sat_mean", "sat_min", "sat_max",
                                          "tau_first", "err_first")] + \
               [("feat", C.c_float * 16), ("wall_ms", C.c_double)] + \
               [(n, C.c_int) for n in ("m", "mD", "rejects", "fevals", "controller", "m_cap")] + \
               [("tau_next", C.c_double)]


class tsr_run_stats(C.Structure):
    _fields_ = [("t", C.c_double), ("steps", C.c_int), ("rejects", C.c_int), ("fevals", C.c_int),
                ("failed", C.c_int), ("wall_s", C.c_double), ("gfunc_ms", C.c_double), ("arnoldi_ms", C.c_double),
                ("host_dense_s", C.c_double), ("sync_wait_s", C.c_double)]


GFUNC_T = C.CFUNCTYPE(C.c_int, _dp, _dp, user_data_struct)

HOST_ELEMENT, HOST_NODE = 0, 1
CFG_PLAIN_SMALL, CFG_PLAIN, CFG_RAGGED, CFG_SCENARIOS, CFG_LARGE = 0, 1, 2, 3, 4
PATH_PLAIN, PATH_DCGS2, PATH_DEFERRED, PATH_JVLIN, PATH_MONOTONE, PATH_BATCHED, PATH_F32BASIS = 0, 1, 2, 4, 8, 15, 16
PATH_GRAPH = 32
KT_NAMES = ("k_micro (full RHS)", "k_micro (linearised FD product)", "k_macro_node", "k_pass_dca", "k_pass_dcb",
            "k_pass")
KT_COUNT = len(KT_NAMES)

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libtsr.so")


def declare(lib):
    """Argument and result types of every exported C function of libtsr.so."""
    P = C.c_void_p
    sig = {
        "Gfunc_hip": (C.c_int, [_dp, _dp, user_data_struct]),
        "exprem_c": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_double, _dp, _dp, _dp, user_data_struct,
                               exprem_options_struct, C.POINTER(exprem_stats_struct)]),
        "phipade": (C.c_int, [_dp, _dp, C.c_double, C.c_int, C.c_int]),
        "tsr_phi1_e1": (C.c_int, [_dp, _dp, C.c_double, C.c_int, C.c_int, C.c_int]),
        "tsr_phik_e1": (C.c_int, [_dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]),
        "weighted_norm": (C.c_double, [_dp, _dp, C.c_double, C.c_double, C.c_int]),
        "check_step": (None, [C.c_double, C.c_double]),
        "check_flag_Gfunc": (C.c_double, [C.c_int, C.c_double]),
        "set_exprem_options_c": (C.c_int, [C.POINTER(exprem_options_struct)]),
        "exprem_stats_init": (None, [C.POINTER(exprem_stats_struct), C.POINTER(exprem_options_struct), _dp]),
        "tsr_problem_config": (P, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong]),
        "tsr_problem_reference": (P, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(soil_struct),
                                      C.POINTER(soil_struct), C.c_double, _ip, C.c_int, C.c_double, C.c_int]),
        "tsr_problem_free": (None, [P]),
        "tsr_problem_view": (C.POINTER(tsr_batch_view), [P]),
        "tsr_problem_user_data": (C.POINTER(user_data_struct), [P]),
        "tsr_problem_N": (C.c_longlong, [P]),
        "tsr_problem_initial": (None, [P, _dp]),
        "tsr_problem_tend": (C.c_double, [P]),
        "tsr_problem_options": (None, [P, C.POINTER(exprem_options_struct)]),
        "tsr_problem_distinct_bytes": (C.c_longlong, [P]),
        "tsr_last_error": (C.c_char_p, []),
        "tsr_release": (C.c_int, [user_data_struct]),
        "tsr_statistics_write": (C.c_int, [C.c_char_p, C.POINTER(tsr_step_record), C.c_int, C.c_int, C.c_int]),
        "tsr_template_structured": (C.c_int, [C.POINTER(micro_mesh_struct)]),
        "tsr_solver_create": (P, [C.POINTER(tsr_batch_view), C.POINTER(exprem_options_struct)]),
        "tsr_solver_free": (None, [P]),
        "tsr_solver_set_state": (C.c_int, [P, _dp, C.c_double]),
        "tsr_solver_get_state": (C.c_int, [P, _dp, _dp]),
        "tsr_solver_stats": (None, [P, C.POINTER(exprem_stats_struct)]),
        "tsr_solver_reset": (None, [P]),
        "tsr_solver_gfunc_host": (C.c_int, [P, _dp, _dp]),
        "tsr_solver_gfunc_dev": (C.c_int, [P, C.c_void_p, C.c_void_p, C.c_int]),
        "tsr_solver_last_gfunc_ms": (C.c_double, [P]),
        "tsr_solver_time_gfunc_parts": (C.c_int, [P, C.c_void_p, C.c_void_p, C.c_int, _dp]),
        "tsr_solver_time_gfunc_mode": (C.c_int, [P, C.c_void_p, C.c_void_p, C.c_int, C.c_int, _dp]),
        "tsr_solver_time_copy": (C.c_int, [P, C.c_int, _dp]),
        "tsr_solver_micro_shape": (None, [P, _ip, _ip]),
        "tsr_solver_ktimer": (C.c_int, [P, C.c_int]),
        "tsr_host_threads": (C.c_int, [C.c_int]),
        "tsr_solver_ktimer_read": (C.c_int, [P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _dp, _dp]),
        "tsr_solver_set_step": (C.c_int, [P, C.c_double]),
        "tsr_solver_graph_stats": (None, [P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
        "tsr_solver_step": (C.c_int, [P, C.c_double, C.POINTER(tsr_step_record)]),
        "tsr_solver_run": (C.c_int, [P, C.c_double, C.c_int, C.POINTER(tsr_run_stats),
                                     C.POINTER(tsr_step_record), C.c_int]),
        "tsr_solver_set_controller": (C.c_int, [P, C.c_int]),
        "tsr_solver_set_path": (C.c_int, [P, C.c_int]),
        "tsr_solver_set_seed": (None, [P, C.c_ulonglong]),
        "tsr_solver_jv": (C.c_int, [P, _dp, C.c_double, C.c_int, _dp]),
        "tsr_solver_jv_noalloc": (C.c_int, [P, _dp, C.c_double, C.c_int, _dp]),
        "tsr_solver_set_mlp": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "tsr_mlp_param_count": (C.c_int, []),
        "tsr_mlp_create": (P, []),
        "tsr_mlp_free": (None, [P]),
        "tsr_mlp_set_params": (C.c_int, [P, _fp]),
        "tsr_mlp_get_params": (C.c_int, [P, _fp]),
        "tsr_mlp_forward": (C.c_int, [P, _fp, C.c_int, _fp]),
        "tsr_mlp_grad": (C.c_int, [P, _fp, _fp, C.c_int, C.c_float, C.c_float, C.c_float, _fp, _fp]),
        "tsr_mlp_adam": (C.c_int, [P, C.c_float]),
        "tsr_mlp_train": (C.c_int, [P, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                    C.c_ulonglong, _fp]),
        "tsr_solver_dev_u": (C.c_void_p, [P]),
        "tsr_solver_dev_scratch": (C.c_void_p, [P, C.c_int]),
        "tsr_solver_N": (C.c_longlong, [P]),
        "tsr_device_sync": (C.c_int, []),
        "tsr_h2d": (C.c_int, [C.c_void_p, _dp, C.c_longlong]),
        "tsr_d2h": (C.c_int, [_dp, C.c_void_p, C.c_longlong]),
    }
    missing = []
    for name, (res, args) in sig.items():
        try:
            f = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        f.restype = res
        f.argtypes = args
    return missing


EXPORTED = None  # filled lazily: names declare() knows about
