"""Native problem plugins on the device: the BiSphere plugin (f1 = sum x^2,
f2 = sum (x - offset)^2, cv = max(f1 - r^2, 0)) evaluated inside NSGA-III and
RVEA runs through the loader's C entry points, checked against the formula
in float64 after initialize, step and reset, on a side stream
(tests/python/test_native_problem.py:233-269 of the reference: rtol 2e-5,
atol 2e-4)."""
import gc
from pathlib import Path

import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402
from cuda_moea.native import NativeProblem, benchmark  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
PLUGIN = Path(__file__).resolve().parent.parent / "cuda-moea_amd" / "lib" / "plugins" / "libbi_sphere.so"


def _problem(**kw):
    opts = dict(library=PLUGIN, name="BiSphere", dimension=7, objectives=2, lower_bounds=-5.0, upper_bounds=5.0)
    opts.update(kw)
    return NativeProblem(**opts)


def _expected(x, offset, radius):
    xd = x.double()
    f = torch.stack((xd.square().sum(1), (xd - offset).square().sum(1)), 1)
    return f, (f[:, 0] - radius * radius).clamp_min(0)


def test_unconstrained_run_and_benchmark():
    p = _problem(parameters={"offset": 2.0})
    algo = cm.NSGA3(problem=p, population_size=32, max_generations=1, print_progress=False)
    print("native BiSphere NSGA-III unconstrained", flush=True)
    result = algo.run()
    torch.testing.assert_close(result.constraints, torch.zeros_like(result.constraints))
    f, _ = _expected(result.variables, 2.0, 3.0)
    torch.testing.assert_close(result.objectives.double(), f, rtol=2e-5, atol=2e-4)
    timing = benchmark(p, result.variables, repeats=3, warmup=1)
    assert timing["wall_ms_per_eval"] > 0 and timing["cuda_ms_per_eval"] > 0


@pytest.mark.parametrize("algorithm", ["NSGA3", "RVEA"])
@pytest.mark.parametrize("dimension", [7, 129])
def test_objectives_constraints_streams_reset_and_lifetime(algorithm, dimension):
    p = _problem(dimension=dimension, parameters={"offset": 1.5, "radius": 2.0, "constrained": True})
    with torch.cuda.stream(torch.cuda.Stream(DEV)):
        algo = getattr(cm, algorithm)(problem=p, population_size=32, max_generations=2, print_progress=False)
        algo.initialize()
        for name, action in (("initialize", lambda: None), ("step", algo.step), ("reset", algo.reset)):
            action()
            print(f"native BiSphere {algorithm} d={dimension} after {name}", flush=True)
            pop = algo.population
            torch.cuda.current_stream(DEV).synchronize()
            # RVEA's dead rows hold NaN variables and sentinel objectives
            valid = torch.isfinite(pop.variables).all(1)
            assert valid.any()
            x = pop.variables[valid]
            f, cv = _expected(x, 1.5, 2.0)
            torch.testing.assert_close(pop.objectives[valid].double(), f, rtol=2e-5, atol=2e-4)
            torch.testing.assert_close(pop.constraints[valid].double(), cv, rtol=2e-5, atol=2e-4)
        result = algo.run(copy=False)
        del algo, p
        gc.collect()
        torch.cuda.synchronize()
        assert result.active_count > 0
        assert torch.isfinite(result.objectives[:result.active_count]).all()


def test_large_batch_evaluation_matches_float64():
    """One evaluation of N = 4096 rows of D = 1024 (the reference's native
    evaluation benchmark shape) against float64: each sum adds 1024 float32
    products, so its rounding error stays within 1024 * 2^-24 relative."""
    p = _problem(dimension=1024, lower_bounds=-1.0, upper_bounds=1.0,
                 parameters={"offset": 0.25, "radius": 18.5, "constrained": True})
    g = torch.Generator(device=DEV).manual_seed(7)
    x = torch.rand(4096, 1024, generator=g, device=DEV) * 2 - 1
    inst = p._instantiate()
    stream = torch.cuda.current_stream(DEV)
    bounds = torch.tensor([-1.0, 1.0] * 1024, device=DEV)
    f = torch.full((2, 4096), float("nan"), device=DEV)
    cv = torch.full((4096,), float("nan"), device=DEV)
    inst.initialize(stream.cuda_stream, 0, bounds.data_ptr())
    from cuda_moea import _lib
    print("native BiSphere N=4096 D=1024 evaluation", flush=True)
    _lib.check(_lib.lib.moea_plugin_evaluate(inst.handle, x.data_ptr(), 4096, 1024, 2, f.data_ptr(),
                                             cv.data_ptr(), 0, 1, stream.cuda_stream))
    stream.synchronize()
    ref_f, ref_cv = _expected(x, 0.25, 18.5)
    tol = 1024 * 2.0 ** -24
    assert bool(((f.t().double() - ref_f).abs() <= tol * ref_f.abs() + 1e-6).all())
    assert bool(((cv.double() - ref_cv).abs() <= tol * ref_f[:, 0] + 1e-6).all())
    assert bool((cv > 0).any()) and bool((cv == 0).any())
    inst.destroy()


def test_native_runs_are_deterministic():
    """The same seeded run twice gives the same bits, for both algorithms."""
    for algorithm in ("NSGA3", "RVEA"):
        print(f"native BiSphere {algorithm} determinism", flush=True)
        runs = [getattr(cm, algorithm)(problem=_problem(dimension=20, parameters={"constrained": True}),
                                       population_size=64, max_generations=5, seed=3,
                                       print_progress=False).run() for _ in range(2)]
        assert torch.equal(runs[0].variables.nan_to_num(7.0), runs[1].variables.nan_to_num(7.0))
        assert torch.equal(runs[0].objectives, runs[1].objectives)
        assert torch.equal(runs[0].constraints, runs[1].constraints)
