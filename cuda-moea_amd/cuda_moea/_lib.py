"""ctypes binding of libmoea_hip.so (C ABI declared in include/moea_hip.h).

The library is the only compute path: if it is missing this module raises at
import time instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int, c_int32, c_size_t, c_uint64, c_void_p
from pathlib import Path

_PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = _PKG_ROOT / "lib" / "libmoea_hip.so"

MOEA_OK = 0
# max_front value of a sort whose front-peeling launch gave up waiting at its
# grid barrier; stats word STAT_DEVICE_ERROR is the sticky flag of the same
DEVICE_FAILED = -2
STAT_DEVICE_ERROR = 6
MATING_TOURNAMENT, MATING_RANDOM, MATING_GIVEN = 0, 1, 2
MUTATION_NONE, MUTATION_POLYNOMIAL = 0, 1
ALGORITHM_NSGA3, ALGORITHM_RVEA = 0, 1
PROBLEM_QUADRATIC, PROBLEM_MLP, PROBLEM_CALLBACK, PROBLEM_DTLZ1, PROBLEM_DTLZ2 = 0, 1, 2, 3, 4
PROBLEM_DTLZ3, PROBLEM_DTLZ4, PROBLEM_DTLZ5, PROBLEM_DTLZ6, PROBLEM_DTLZ7, PROBLEM_CONVEX_DTLZ2 = 5, 6, 7, 8, 9, 10
# native DTLZ problem kinds by name (moea_eval_dtlz numbers the same problems 1-15)
DTLZ_KINDS = {"DTLZ1": PROBLEM_DTLZ1, "DTLZ2": PROBLEM_DTLZ2, "DTLZ3": PROBLEM_DTLZ3, "DTLZ4": PROBLEM_DTLZ4,
              "DTLZ5": PROBLEM_DTLZ5, "DTLZ6": PROBLEM_DTLZ6, "DTLZ7": PROBLEM_DTLZ7,
              "ConvexDTLZ2": PROBLEM_CONVEX_DTLZ2, "C1DTLZ1": 11, "C1DTLZ3": 12, "C2DTLZ2": 13,
              "C2ConvexDTLZ2": 14, "C3DTLZ1": 15, "C3DTLZ4": 16, "CSDP": 17}
MAX_OBJECTIVES = 10


class Draws(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in (
        "mate_a", "mate_b", "sbx_u", "sbx_p", "sbx_sign", "sbx_copy", "pm_gate", "pm_delta",
        "niche_a", "niche_b")]


class PopulationC(ctypes.Structure):
    _fields_ = [("x", c_void_p), ("f", c_void_p), ("cv", c_void_p)]


class CvQuantization(ctypes.Structure):
    _fields_ = [("bins", c_int), ("clip_upper", c_float), ("log_alpha", c_float),
                ("feasibility_epsilon", c_float)]


class Niching(ctypes.Structure):
    _fields_ = [("mode", c_int), ("lds_cap", c_int)]


class NSGA3SelectArgs(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("m", c_int), ("d", c_int), ("k", c_int),
        ("parents", PopulationC), ("offspring", PopulationC), ("next", PopulationC),
        ("next_fronts", c_void_p), ("directions", c_void_p), ("ideal", c_void_p),
        ("draws", POINTER(Draws)), ("seed", c_uint64), ("generation", c_int),
        ("directions_ready", c_int),
        ("stats", c_void_p), ("workspace", c_void_p), ("workspace_bytes", c_size_t),
        ("stream", c_void_p), ("cv_quant", CvQuantization),
        ("out_extremes", c_void_p), ("out_intercepts", c_void_p), ("niching", Niching),
    ]


class ReproduceArgs(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("d", c_int), ("active", c_int), ("active_device", c_void_p),
        ("mating", c_int), ("parents_x", c_void_p), ("parents_rank", c_void_p),
        ("parents_cv", c_void_p), ("bounds", c_void_p), ("sbx_eta", c_float),
        ("sbx_prob", c_float), ("sbx_copy_prob", c_float), ("mutation", c_int),
        ("pm_eta", c_float), ("pm_prob", c_float), ("draws", POINTER(Draws)),
        ("seed", c_uint64), ("generation", c_int), ("offspring_x", c_void_p),
        ("parent_indices", c_void_p), ("stream", c_void_p),
    ]


class RVEASelectArgs(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("m", c_int), ("d", c_int), ("k", c_int), ("active", c_int),
        ("active_device", c_void_p),
        ("parents", PopulationC), ("offspring", PopulationC), ("next", PopulationC),
        ("directions", c_void_p), ("gamma", c_void_p), ("penalty", c_float),
        ("out_active", c_void_p), ("workspace", c_void_p), ("workspace_bytes", c_size_t),
        ("stream", c_void_p),
    ]


EVALUATE_FN = ctypes.CFUNCTYPE(c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                               c_void_p, c_int, c_int, c_void_p)


_POP_P = ctypes.POINTER(PopulationC)
PREPARE_PARENT_FN = ctypes.CFUNCTYPE(c_int, c_void_p, _POP_P, c_int, c_int, c_int, c_int, c_int, c_void_p)


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("kind", c_int), ("centers", c_void_p), ("w1", c_void_p), ("w2", c_void_p),
        ("hidden", c_int), ("hwork", c_void_p), ("callback", EVALUATE_FN), ("user", c_void_p),
        ("cv_activation_ratio", c_float), ("prepare_parent", PREPARE_PARENT_FN),
    ]
MATE_FN = ctypes.CFUNCTYPE(c_int, c_void_p, _POP_P, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                           c_int, c_int, c_void_p)
CROSSOVER_FN = ctypes.CFUNCTYPE(c_int, c_void_p, _POP_P, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                                c_int, c_void_p)
MUTATE_FN = ctypes.CFUNCTYPE(c_int, c_void_p, _POP_P, c_int, c_int, c_int, c_int, c_int, c_void_p)
SELECT_FN = ctypes.CFUNCTYPE(c_int, c_void_p, _POP_P, _POP_P, _POP_P, c_int, c_int, c_int, c_int, c_int,
                             c_void_p)
PREPARE_FN = ctypes.CFUNCTYPE(c_int, c_void_p, _POP_P, c_int, c_int, c_int, c_int, c_int, c_void_p)
DIRECTIONS_FN = ctypes.CFUNCTYPE(c_int, c_void_p, _POP_P, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                                 c_int, c_int, c_void_p)


class StrategyHooks(ctypes.Structure):
    _fields_ = [
        ("user", c_void_p), ("mate", MATE_FN), ("crossover", CROSSOVER_FN), ("mutate", MUTATE_FN),
        ("update_directions", DIRECTIONS_FN), ("select", SELECT_FN), ("prepare", PREPARE_FN),
    ]


class RunArgs(ctypes.Structure):
    _fields_ = [
        ("algorithm", c_int), ("n", c_int), ("m", c_int), ("d", c_int), ("k", c_int),
        ("max_generations", c_int), ("pops", PopulationC * 3), ("current", c_int),
        ("fronts", c_void_p * 2), ("fronts_current", c_int), ("parent_indices", c_void_p),
        ("bounds", c_void_p), ("directions", c_void_p), ("directions0", c_void_p),
        ("gamma", c_void_p), ("ideal", c_void_p), ("active_device", c_void_p),
        ("adapt_scratch", c_void_p), ("stats", c_void_p), ("stats_init", c_void_p),
        ("mating", c_int), ("mutation", c_int), ("sbx_eta", c_float * 2), ("sbx_prob", c_float),
        ("sbx_copy_prob", c_float), ("pm_eta", c_float * 2), ("pm_prob", c_float),
        ("rvea_alpha", c_float), ("rvea_frequency", c_float), ("seed", c_uint64),
        ("initial_x_given", c_int), ("problem", ProblemDesc), ("generation", c_int),
        ("directions_ready", c_int),
        ("workspace", c_void_p), ("workspace_bytes", c_size_t), ("stream", c_void_p),
        ("hooks", ctypes.POINTER(StrategyHooks)), ("cv_quant", CvQuantization), ("niching", Niching),
    ]


class BatchProblem(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("m", c_int), ("d", c_int), ("k", c_int), ("max_generations", c_int),
        ("seed", c_uint64), ("problem", c_int), ("centers", c_void_p), ("w1", c_void_p),
        ("w2", c_void_p), ("hidden", c_int), ("bounds", c_void_p), ("directions", c_void_p),
        ("sbx_eta", c_float * 2), ("sbx_prob", c_float), ("sbx_copy_prob", c_float),
        ("pm_eta", c_float * 2), ("pm_prob", c_float), ("niching", Niching),
    ]


class ServeConfig(ctypes.Structure):
    _fields_ = [("max_instances", c_int), ("max_rows", ctypes.c_longlong), ("niching", Niching)]


PARAM_INT, PARAM_FLOAT, PARAM_BOOL, PARAM_STR, PARAM_FLOAT_LIST = 0, 1, 2, 3, 4
PARAM_TYPES = ("int", "float", "bool", "str", "float_list")


class ParamValue(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("type", c_int), ("list_length", c_int),
                ("int_value", ctypes.c_longlong), ("float_value", ctypes.c_double),
                ("str_value", ctypes.c_char_p), ("list", POINTER(ctypes.c_double))]


# name -> (restype, argtypes); every symbol include/moea_hip.h declares
SIGNATURES = {
    "moea_abi_version": (c_int, []),
    "moea_last_error": (ctypes.c_char_p, []),
    "moea_nds_workspace_bytes": (c_size_t, [c_int]),
    "moea_nds_rank": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                              c_void_p, c_size_t, c_void_p]),
    "moea_nds_rank_static": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                     c_void_p]),
    "moea_nds_static_max_n": (c_int, []),
    "moea_debug_set_spin_limit": (None, [ctypes.c_uint32]),
    "moea_das_dennis": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p]),
    "moea_nsga3_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "moea_nsga3_select": (c_int, [POINTER(NSGA3SelectArgs)]),
    "moea_ideal_point": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "moea_associate_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "moea_associate": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                               c_int, c_void_p, c_size_t, c_void_p]),
    "moea_reproduce": (c_int, [POINTER(ReproduceArgs)]),
    "moea_quantize_cv": (c_int, [c_void_p, c_void_p, c_int, POINTER(CvQuantization), c_void_p]),
    "moea_annealed": (c_float, [c_float, c_float, c_int, c_int]),
    "moea_init_population": (c_int, [c_void_p, c_int, c_int, c_void_p, c_uint64, c_void_p]),
    "moea_eval_quadratic": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_int,
                                    c_void_p, c_void_p]),
    "moea_eval_dtlz": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "moea_eval_mlp": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                              c_int, c_void_p, c_void_p, c_void_p]),
    "moea_rvea_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "moea_rvea_select": (c_int, [POINTER(RVEASelectArgs)]),
    "moea_rvea_penalty": (c_float, [c_int, c_int, c_int, c_float]),
    "moea_rvea_gamma": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "moea_rvea_adapt": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_void_p,
                                c_void_p, c_void_p]),
    "moea_run_workspace_bytes": (c_size_t, [c_int, c_int, c_int, c_int]),
    "moea_run_init": (c_int, [POINTER(RunArgs)]),
    "moea_run_step": (c_int, [POINTER(RunArgs), c_int]),
    "moea_batch_workspace_bytes": (c_size_t, [POINTER(BatchProblem), c_int]),
    "moea_batch_create": (c_int, [POINTER(BatchProblem), c_int, c_void_p, c_size_t, c_void_p,
                                  POINTER(c_void_p)]),
    "moea_batch_init": (c_int, [c_void_p]),
    "moea_batch_step": (c_int, [c_void_p, c_int]),
    "moea_batch_generation": (c_int, [c_void_p]),
    "moea_batch_view": (c_int, [c_void_p, c_int, POINTER(PopulationC), POINTER(c_void_p)]),
    "moea_batch_check": (c_int, [c_void_p]),
    "moea_batch_set_graph": (c_int, [c_void_p, c_int]),
    "moea_batch_graph_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "moea_batch_peel_plan": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int)]),
    "moea_batch_set_peel_plan": (c_int, [c_void_p, c_int, c_int]),
    "moea_batch_autotune": (c_int, [c_void_p, c_int, POINTER(c_int), POINTER(ctypes.c_double), c_int]),
    "moea_batch_destroy": (None, [c_void_p]),
    "moea_serve_reserved_bytes": (c_size_t, [POINTER(ServeConfig)]),
    "moea_serve_instance_bytes": (c_size_t, [POINTER(BatchProblem)]),
    "moea_serve_create": (c_int, [POINTER(ServeConfig), c_void_p, c_size_t, c_void_p, POINTER(c_void_p)]),
    "moea_serve_admit": (c_int, [c_void_p, POINTER(BatchProblem), POINTER(c_int)]),
    "moea_serve_step": (c_int, [c_void_p, c_void_p, c_int, POINTER(c_int)]),
    "moea_serve_generation": (c_int, [c_void_p, c_int]),
    "moea_serve_view": (c_int, [c_void_p, c_int, POINTER(PopulationC), POINTER(c_void_p)]),
    "moea_serve_release": (c_int, [c_void_p, c_int]),
    "moea_serve_check": (c_int, [c_void_p]),
    "moea_serve_set_graph": (c_int, [c_void_p, c_int]),
    "moea_serve_destroy": (None, [c_void_p]),
    "moea_native_build_id": (ctypes.c_char_p, []),
    "moea_plugin_open": (c_int, [ctypes.c_char_p, POINTER(c_void_p)]),
    "moea_plugin_close": (None, [c_void_p]),
    "moea_plugin_name": (ctypes.c_char_p, [c_void_p]),
    "moea_plugin_build_id": (ctypes.c_char_p, [c_void_p]),
    "moea_plugin_param_count": (c_int, [c_void_p]),
    "moea_plugin_param": (c_int, [c_void_p, c_int, POINTER(ParamValue), POINTER(ctypes.c_char_p)]),
    "moea_plugin_create": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, POINTER(ParamValue), c_int,
                                   POINTER(c_void_p)]),
    "moea_plugin_problem_constraints": (c_int, [c_void_p]),
    "moea_plugin_problem_initialize": (c_int, [c_void_p, c_void_p, c_uint64, c_void_p]),
    "moea_plugin_problem_destroy": (None, [c_void_p]),
    "moea_plugin_evaluate": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                     c_void_p]),
    "moea_plugin_prepare_parent": (c_int, [c_void_p, _POP_P, c_int, c_int, c_int, c_int, c_int, c_void_p]),
    "moea_plugin_reset": (c_int, [c_void_p]),
}


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"libmoea_hip.so not found at {LIB_PATH}; build it with "
            "`make -C cuda-moea_amd` (hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class MoeaError(RuntimeError):
    """A libmoea_hip call failed on the device side."""


def check(status: int) -> None:
    """Raise the exception type the reference raises for the same failure:
    invalid configurations -> ValueError, device problems -> RuntimeError."""
    if status == MOEA_OK:
        return
    msg = (lib.moea_last_error() or b"").decode(errors="replace")
    if status in (-1, -4):
        raise ValueError(msg)
    raise MoeaError(f"libmoea_hip error {status}: {msg}")


# optional per-stage timing (profiling builds nothing extra; off by default)
for _name, _res, _args in (("moea_profile_enable", c_int, [c_int]),
                           ("moea_profile_read", c_int, [c_void_p, c_void_p, c_int]),
                           ("moea_profile_stage_name", ctypes.c_char_p, [c_int])):
    _fn = getattr(lib, _name)
    _fn.restype, _fn.argtypes = _res, _args
    SIGNATURES[_name] = (_res, _args)


def profile_enable(on: bool) -> None:
    check(lib.moea_profile_enable(1 if on else 0))


def profile_read() -> dict:
    """Per-stage {name: (total_ms, launches)} since the last read (synchronizes)."""
    import numpy as _np
    ms = _np.zeros(32, _np.float64)
    cnt = _np.zeros(32, _np.int64)
    n = lib.moea_profile_read(ms.ctypes.data_as(c_void_p), cnt.ctypes.data_as(c_void_p), 32)
    return {lib.moea_profile_stage_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}
