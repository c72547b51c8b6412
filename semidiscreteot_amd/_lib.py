"""ctypes binding of libsdot.so (include/sdot.h).

The product path has no fallback: if the gfx950 library is missing, importing
anything that needs it raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsdot.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u64p = C.POINTER(C.c_uint64)

OK = 0
ERR_NAMES = {-1: "SDOT_ERR_INVALID", -2: "SDOT_ERR_STATE", -3: "SDOT_ERR_HIP",
             -4: "SDOT_ERR_NOMEM", -5: "SDOT_ERR_PARSE", -6: "SDOT_ERR_UNSUPPORTED"}
SOLVE_STATUS = {0: "converged", 1: "max_iterations", 2: "linesearch_fail", 3: "nonfinite"}


class SdotError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("max_iterations", C.c_uint32), ("tolerance", C.c_double),
                ("solver_control_type", C.c_int32), ("epsilon", C.c_double),
                ("tau", C.c_double), ("distance_threshold_type", C.c_int32),
                ("verbose_output", C.c_int32), ("solver_type", C.c_int32),
                ("quadrature_order", C.c_uint32), ("nb_points", C.c_uint32),
                ("number_of_threads", C.c_uint32), ("use_epsilon_scaling", C.c_int32),
                ("epsilon_scaling_factor", C.c_double), ("epsilon_scaling_steps", C.c_uint32),
                ("use_log_sum_exp_trick", C.c_int32), ("precision", C.c_int32),
                ("line_search", C.c_int32), ("preconditioner", C.c_int32),
                ("prune_bits", C.c_double)]


class MeshParams(C.Structure):
    _fields_ = [("n_refinements", C.c_uint32), ("grid_generator_function", C.c_char * 64),
                ("grid_generator_arguments", C.c_char * 256), ("use_tetrahedral_mesh", C.c_int32),
                ("use_custom_density", C.c_int32), ("density_file_path", C.c_char * 256),
                ("density_file_format", C.c_char * 16), ("density_field_name", C.c_char * 64)]


class MultilevelParams(C.Structure):
    _fields_ = [("source_enabled", C.c_int32), ("source_min_vertices", C.c_int32),
                ("source_max_vertices", C.c_int32), ("source_hierarchy_dir", C.c_char * 256),
                ("target_enabled", C.c_int32), ("target_min_points", C.c_int32),
                ("target_max_points", C.c_int32), ("target_hierarchy_dir", C.c_char * 256),
                ("use_softmax_potential_transfer", C.c_int32),
                ("use_python_clustering", C.c_int32), ("python_script_name", C.c_char * 128),
                ("output_prefix", C.c_char * 256)]


MAX_DENSITY_INDICES = 256


class PrmFile(C.Structure):
    _fields_ = [("solver", Params), ("selected_task", C.c_char * 64), ("io_coding", C.c_int32),
                ("source", MeshParams), ("target", MeshParams), ("multilevel", MultilevelParams),
                ("power_diagram_implementation", C.c_char * 16),
                ("transport_map_truncation_radius", C.c_double),
                ("n_density_indices", C.c_int32),
                ("density_indices", C.c_uint32 * MAX_DENSITY_INDICES),
                ("density_potential_folder", C.c_char * 256),
                ("density_truncation_radius", C.c_double),
                ("lloyd_max_iterations", C.c_uint32), ("lloyd_relative_tolerance", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("evaluations", C.c_int32),
                ("functional", C.c_double), ("check_value", C.c_double), ("seconds", C.c_double)]


class State(C.Structure):
    _fields_ = [("last_functional", C.c_double), ("C_global", C.c_double),
                ("last_distance_threshold", C.c_double), ("covering_radius", C.c_double),
                ("min_target_density", C.c_double), ("last_iteration_count", C.c_int32),
                ("converged", C.c_int32), ("last_pair_count", C.c_int64)]


class Profile(C.Structure):
    _fields_ = [("evaluations", C.c_int64), ("ms_lse_max", C.c_double),
                ("ms_lse_sum", C.c_double), ("ms_qfin", C.c_double), ("ms_grad", C.c_double),
                ("ms_total", C.c_double), ("hessians", C.c_int64), ("ms_hessian", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_slots", C.c_int32),
                ("hbm_budget_bytes", C.c_int64), ("use_graph", C.c_int32),
                ("pass_targets", C.c_int32), ("grad_qpoints", C.c_int32),
                ("polish_factor", C.c_double)]


class Request(C.Structure):
    _fields_ = [("n", C.c_int64), ("y", _dp), ("nu", _dp), ("psi0", _dp),
                ("epsilon", C.c_double), ("tolerance", C.c_double),
                ("solver_control_type", C.c_int32), ("max_iterations", C.c_uint32),
                ("radius", C.c_double), ("line_search", C.c_int32),
                ("epsilon_scaling_factor", C.c_double), ("epsilon_scaling_steps", C.c_uint32),
                ("preconditioner", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("iterations", C.c_int32),
                ("evaluations", C.c_int32), ("functional", C.c_double),
                ("check_value", C.c_double), ("rel_marginal_error", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


# name -> (restype, argtypes); every function declared in include/sdot.h
SIGNATURES = {
    "sdot_version": (C.c_char_p, []),
    "sdot_last_error": (C.c_char_p, []),
    "sdot_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "sdot_ctx_destroy": (C.c_int, [C.c_void_p]),
    "sdot_params_default": (C.c_int, [C.POINTER(Params)]),
    "sdot_params_parse_prm": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
    "sdot_params_load_prm": (C.c_int, [C.c_char_p, C.POINTER(Params)]),
    "sdot_prm_file_default": (C.c_int, [C.POINTER(PrmFile)]),
    "sdot_prm_file_parse": (C.c_int, [C.c_char_p, C.POINTER(PrmFile)]),
    "sdot_prm_file_load": (C.c_int, [C.c_char_p, C.POINTER(PrmFile)]),
    "sdot_set_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
    "sdot_set_source": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int, _dp, _dp, _dp]),
    "sdot_set_target": (C.c_int, [C.c_void_p, C.c_int64, _dp, _dp]),
    "sdot_target_layout": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_int64), _dp]),
    "sdot_eval": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp]),
    "sdot_eval_R": (C.c_int, [C.c_void_p, C.c_double, _dp, _dp, _dp, _dp]),
    "sdot_eval_device": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sdot_eval_device_noalloc": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "sdot_synchronize": (C.c_int, [C.c_void_p]),
    "sdot_hessian": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int64]),
    "sdot_barycenters": (C.c_int, [C.c_void_p, _dp, _dp]),
    "sdot_barycenters_R": (C.c_int, [C.c_void_p, C.c_double, _dp, _dp]),
    "sdot_softmax_refinement": (C.c_int, [C.c_void_p, C.c_int64, _dp, C.c_int64, _dp, _dp, _dp,
                                          C.c_int64, _i64p, _i64p, C.c_double, C.c_double,
                                          C.c_int, _dp]),
    "sdot_conditional_density": (C.c_int, [C.c_void_p, _dp, C.c_double, C.c_int64, _dp, _dp,
                                           C.c_int64, _i64p, _dp]),
    "sdot_transport_map": (C.c_int, [C.c_void_p, C.c_int, _dp, C.c_double, C.c_double,
                                     C.c_int64, _dp, _dp]),
    "sdot_solve": (C.c_int, [C.c_void_p, _dp, C.POINTER(SolveInfo)]),
    "sdot_covering_radius": (C.c_int, [C.c_void_p, _dp]),
    "sdot_distance_threshold": (C.c_int, [C.c_void_p, _dp, _dp]),
    "sdot_set_distance_threshold": (C.c_int, [C.c_void_p, C.c_double]),
    "sdot_candidate_stats": (C.c_int, [C.c_void_p, C.c_double, _i64p, _u64p, _i64p]),
    "sdot_q_log_partition": (C.c_int, [C.c_void_p, C.c_double, _dp, _dp]),
    "sdot_get_state": (C.c_int, [C.c_void_p, C.POINTER(State)]),
    "sdot_set_state": (C.c_int, [C.c_void_p, C.POINTER(State)]),
    "sdot_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "sdot_profile_get": (C.c_int, [C.c_void_p, C.POINTER(Profile)]),
    "sdot_server_config_default": (C.c_int, [C.POINTER(ServerConfig)]),
    "sdot_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "sdot_server_destroy": (C.c_int, [C.c_void_p]),
    "sdot_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), _i64p]),
    "sdot_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "sdot_server_poll": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int32),
                                   C.POINTER(Result), _dp]),
    "sdot_server_clock": (C.c_int, [C.c_void_p, _dp]),
    "sdot_server_request_bytes": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, _i64p]),
    "sdot_server_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32)]),
    "sdot_mesh_structured": (C.c_int, [C.c_int, C.POINTER(C.c_int), _dp, _dp, C.c_int, _dp,
                                       C.c_int, _dp, _dp, _dp]),
    "sdot_mesh_simplex": (C.c_int, [C.c_int, C.c_int64, _dp, C.c_int64, _i64p, C.c_int, _dp,
                                    C.c_int, _dp, _dp, _dp]),
    "sdot_mesh_cells": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int64, _dp, C.c_int64, _i64p,
                                  C.c_int, _dp, C.c_int, _dp, _dp, _dp]),
}

_lib = None


def load():
    """Load libsdot.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m semidiscreteot_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = load().sdot_last_error().decode(errors="replace")
        raise SdotError(rc, msg)
    return rc
