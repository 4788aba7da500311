"""The C-ABI library loads and exports every function include/moea_hip.h
declares (no compute calls; CPU only)."""
import ctypes
import re
import subprocess

import pytest
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "moea_hip.h"
LIB = ROOT / "cuda-moea_amd" / "lib" / "libmoea_hip.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(moea_[a-z0-9_]+)\s*\(", text)) - {"moea_evaluate_fn"})


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("moea_nds_rank", "moea_nsga3_select", "moea_reproduce", "moea_rvea_select",
                 "moea_eval_quadratic", "moea_run_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        subprocess.run(["make", "-s", "-j8", "-C", str(ROOT / "cuda-moea_amd")], check=True)
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import cuda_moea._lib as L
    assert set(declared_functions()) <= set(L.SIGNATURES)


def test_host_side_calls():
    import cuda_moea._lib as L
    assert L.lib.moea_abi_version() == 4
    assert L.lib.moea_das_dennis(92, 3, 0, 0, None, None) == 91
    assert L.lib.moea_das_dennis(32768, 3, 0, 0, None, None) == 32640
    assert abs(L.lib.moea_annealed(30.0, 10.0, 50, 100) - 20.0) < 1e-6
    assert abs(L.lib.moea_rvea_penalty(5, 50, 100, 2.0) - 1.25) < 1e-6
    assert L.lib.moea_das_dennis(92, 11, 0, 0, None, None) < 0
    assert b"objective" in L.lib.moea_last_error()


STRUCTS = {"Draws": "moea_draws", "PopulationC": "moea_population", "NSGA3SelectArgs": "moea_nsga3_select_args",
           "ReproduceArgs": "moea_reproduce_args", "RVEASelectArgs": "moea_rvea_select_args",
           "ProblemDesc": "moea_problem_desc", "RunArgs": "moea_run_args", "BatchProblem": "moea_batch_problem",
           "ServeConfig": "moea_serve_config", "StrategyHooks": "moea_strategy_hooks",
           "CvQuantization": "moea_cv_quantization", "Niching": "moea_niching"}


def test_ctypes_structs_match_header_layout(tmp_path):
    """Every ctypes mirror has the header struct's size and field offsets
    (gcc on the header itself)."""
    import shutil
    import pytest
    import cuda_moea._lib as L
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "moea_hip.h"', "int main(void) {"]
    expect = []
    for py, c in STRUCTS.items():
        cls = getattr(L, py)
        lines.append(f'  printf("%zu\\n", sizeof({c}));')
        expect.append(ctypes.sizeof(cls))
        for f in cls._fields_:
            lines.append(f'  printf("%zu\\n", offsetof({c}, {f[0]}));')
            expect.append(getattr(cls, f[0]).offset)
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(HEADER.parent), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expect


def test_das_dennis_matches_oracle():
    import numpy as np
    import oracle as O
    from cuda_moea import kernels as K
    for n, m, p, k in [(92, 3, 0, 0), (210, 5, 0, 0), (16384, 5, 0, 2), (100, 8, 0, 0),
                       (0, 8, 3, 1), (0, 3, 12, 1), (32768, 3, 0, 0)]:
        np.testing.assert_array_equal(K.das_dennis(n, m, p, kind=k), O.das_dennis(n, m, p, kind=k))


def test_api_validation_errors():
    import pytest
    import cuda_moea as cm
    with pytest.raises(ValueError):
        cm.Algorithm("MOEAD")
    with pytest.raises(ValueError):
        cm.SBX(eta_initial=-1.0).validate()
    with pytest.raises(ValueError):
        cm.NSGA3(crossover=cm.SBX(probability=2.0))
    with pytest.raises(ValueError):
        cm.RVEA(environment_selector=cm.RVEAEnvironmentSelector(alpha=0.0))
    # CV quantization parameters as validate_cv_quant_params checks them (cv_quantization.cu:93-112)
    cm.NSGA3(environment_selector=cm.NSGA3EnvironmentSelector(cv_bins=4, cv_clip_upper=2.0, cv_log_alpha=3.0))
    for bad in (dict(cv_clip_upper=0.0), dict(cv_clip_upper=1.0, cv_log_alpha=-1.0),
                dict(cv_clip_upper=float("inf")), dict(cv_clip_upper=1.0, feasibility_epsilon=-0.1)):
        with pytest.raises(ValueError, match="CV quantization"):
            cm.NSGA3(environment_selector=cm.NSGA3EnvironmentSelector(cv_bins=4, **bad))
    # every strategy takes a Python object (adapters.cpp:196-434)
    alg = cm.NSGA3(mating=cm.PythonMating(), crossover=cm.PythonCrossover(), mutation=cm.PythonMutation(),
                   reference_directions=cm.PythonReferenceDirections(),
                   environment_selector=cm.PythonEnvironmentSelector())
    assert not alg.initialized
    # dtlz_problem.cu:63-77: a finite ratio, clamped to [0, 1]; constrained variants declare one constraint
    assert cm.C1DTLZ1(12, 3).constraints == 1 and cm.DTLZ2(12, 3).constraints == 0
    assert cm.C3DTLZ4(12, 3, constraint_activation_ratio=-1.0).constraints == 1
    with pytest.raises(ValueError, match="finite"):
        cm.C1DTLZ1(12, 3, constraint_activation_ratio=float("nan"))
    assert cm.C1DTLZ1(12, 3, constraint_activation_ratio=0.5).constraint_activation_ratio == 0.5
    assert cm.C1DTLZ1(12, 3, constraint_activation_ratio=7.0).constraint_activation_ratio == 1.0
    assert cm.CSDP(7, 3).constraints == 1
    b = cm.AlgorithmBuilder("NSGA3").population_size(92).max_generations(10).problem(
        cm.QuadraticCenters(12, 3))
    alg = b.build()
    assert isinstance(alg, cm.Algorithm) and alg.population_size == 92 and not alg.initialized


def test_serve_admission_host_logic():
    """Slot, row and arena limits of moea_serve_admit (host bookkeeping only:
    the workspace is a host buffer that no device call touches)."""
    import cuda_moea._lib as L
    lib = L.lib
    cfg = L.ServeConfig(max_instances=3, max_rows=1000)
    reserved = lib.moea_serve_reserved_bytes(ctypes.byref(cfg))
    assert reserved > 0 and reserved % 256 == 0
    assert lib.moea_serve_reserved_bytes(ctypes.byref(L.ServeConfig(0, 10))) == 0
    dummy = (ctypes.c_float * 8)()
    p = L.BatchProblem()
    p.n, p.m, p.d, p.k, p.max_generations, p.seed = 100, 3, 20, 91, 5, 1
    p.problem, p.centers = L.PROBLEM_QUADRATIC, ctypes.addressof(dummy)
    p.bounds = p.directions = ctypes.addressof(dummy)
    one = lib.moea_serve_instance_bytes(ctypes.byref(p))
    assert one > 100 * 20 * 4 * 3 and one % 256 == 0
    # arena of two and a half instances: two fit, the third is refused for space
    buf = ctypes.create_string_buffer(reserved + 2 * one + one // 2 + 256)
    srv = ctypes.c_void_p()
    base = (ctypes.addressof(buf) + 255) // 256 * 256
    size = ctypes.sizeof(buf) - (base - ctypes.addressof(buf))
    assert lib.moea_serve_create(ctypes.byref(cfg), base, size, None, ctypes.byref(srv)) == 0
    slot = ctypes.c_int(-1)
    assert lib.moea_serve_admit(srv, ctypes.byref(p), ctypes.byref(slot)) == 0 and slot.value == 0
    assert lib.moea_serve_admit(srv, ctypes.byref(p), ctypes.byref(slot)) == 0 and slot.value == 1
    assert lib.moea_serve_admit(srv, ctypes.byref(p), ctypes.byref(slot)) == -3
    assert b"budget" in lib.moea_last_error()
    assert lib.moea_serve_generation(srv, 1) == 0
    assert lib.moea_serve_generation(srv, 2) < 0
    # row capacity: 2n of a new run beyond max_rows
    q = L.BatchProblem.from_buffer_copy(p)
    q.n = 500
    assert lib.moea_serve_admit(srv, ctypes.byref(q), ctypes.byref(slot)) == -3
    # invalid problems
    q.n = 101
    assert lib.moea_serve_admit(srv, ctypes.byref(q), ctypes.byref(slot)) == -1
    q.n, q.problem = 100, L.PROBLEM_MLP
    assert lib.moea_serve_admit(srv, ctypes.byref(q), ctypes.byref(slot)) == -1
    lib.moea_serve_destroy(srv)


def test_reference_api_constructions():
    """The construction cases of the reference's tests/python/test_api.py:46-110:
    every built-in problem, every built-in strategy combination and every
    Python strategy class construct without a GPU."""
    import torch
    import cuda_moea as cm

    class Sphere(cm.PythonProblem):
        def __init__(self):
            super().__init__(4, 2)

        def evaluate(self, variables, context):
            return torch.stack((variables.square().sum(1), (variables - 1).square().sum(1)), 1)

    class Selector(cm.PythonEnvironmentSelector):
        def select(self, parents, offspring, context):
            return {"indices": torch.arange(parents["variables"].shape[0], device=parents["variables"].device)}

    crossover = cm.SBX(variable_copy_probability=0.5)
    assert crossover._spec()["variable_copy_probability"] == 0.5
    assert not cm.NSGA3(population_size=16, max_generations=1, crossover=crossover, enable_warmup=False).initialized
    assert not cm.NSGA3(population_size=16, max_generations=1, problem=Sphere(),
                        initial_population=torch.zeros((16, 4)), enable_warmup=False).initialized
    for name in ["DTLZ1", "DTLZ2", "DTLZ3", "DTLZ4", "DTLZ5", "DTLZ6", "DTLZ7", "ConvexDTLZ2", "C1DTLZ1",
                 "C1DTLZ3", "C2DTLZ2", "C2ConvexDTLZ2", "C3DTLZ1", "C3DTLZ4", "CSDP"]:
        assert not cm.NSGA3(population_size=16, max_generations=1, problem=getattr(cm, name)(8, 3),
                            enable_warmup=False).initialized
    for options in [dict(mating=cm.RandomMating(), crossover=cm.SBX(), mutation=cm.NoMutation(),
                         reference_directions=cm.DasDennisDirections(2),
                         environment_selector=cm.NSGA3EnvironmentSelector()),
                    dict(reference_directions=cm.UserDefinedDirections(torch.eye(3)))]:
        assert not cm.Algorithm("NSGA3", population_size=16, max_generations=1, enable_warmup=False,
                                **options).initialized
    assert not cm.NSGA3(population_size=16, max_generations=1, problem=Sphere(), mating=cm.PythonMating(),
                        crossover=cm.PythonCrossover(), mutation=cm.PythonMutation(),
                        reference_directions=cm.PythonReferenceDirections(), environment_selector=Selector(),
                        enable_warmup=False).initialized
    assert cm.AlgorithmBuilder("RVEA").population_size(16).max_generations(2).problem(cm.DTLZ2()).build().generation == 0


def test_user_directions_are_count_by_objectives():
    """UserDefinedDirections takes a (count, objective_count) tensor and
    stores it objective-major (adapters.cpp:557-562 of the reference), with
    each direction normalised; a wrong objective count is refused."""
    import numpy as np
    import torch
    import cuda_moea as cm
    vals = torch.tensor([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 3.0], [1.0, 1.0, 0.0]])
    alg = cm.NSGA3(population_size=16, max_generations=1, problem=cm.DTLZ2(8, 3),
                   reference_directions=cm.UserDefinedDirections(vals))
    v = alg._directions_host(16, 3)
    assert v.shape == (3, 4)
    np.testing.assert_allclose(v[:, :3], np.eye(3), atol=0)
    np.testing.assert_allclose(v[:, 3], [2 ** -0.5, 2 ** -0.5, 0.0], rtol=1e-7)
    raw = cm.NSGA3(population_size=16, max_generations=1, problem=cm.DTLZ2(8, 3),
                   reference_directions=cm.UserDefinedDirections(vals, normalize=False))._directions_host(16, 3)
    np.testing.assert_array_equal(raw, vals.numpy().T)
    with pytest.raises(ValueError, match="objective count"):
        cm.NSGA3(population_size=16, max_generations=1, problem=cm.DTLZ2(8, 3),
                 reference_directions=cm.UserDefinedDirections(vals.T.contiguous()))._directions_host(16, 3)
