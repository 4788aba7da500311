"""End-to-end runs through the Python API / native loop against the oracle's
full-run restatement (same in-kernel draw generator), plus API behaviour."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
import cuda_moea as cm  # noqa: E402

pytestmark = pytest.mark.gpu


def _config0(gens, seed=0, n=92, d=12, m=3):
    prob = cm.QuadraticCenters(d, m, seed=seed)
    alg = cm.NSGA3(population_size=n, max_generations=gens, problem=prob, seed=seed,
                   print_progress=False)
    return alg, prob


@pytest.mark.parametrize("gens,seed", [(1, 0), (20, 0), (200, 0), (50, 3)])
def test_nsga3_run_matches_oracle(gens, seed):
    """Config 0 (N=92, M=3, D=12, H=12) trajectory, bit-exact."""
    alg, prob = _config0(gens, seed)
    print(f"nsga3 run gens={gens} seed={seed}", flush=True)
    res = alg.run()
    bounds = np.tile(np.array([0.0, 1.0], np.float32), 12)
    x, f, cv, fr, _ = O.nsga3_run(92, 3, 12, gens, seed, bounds, prob.centers.numpy())
    np.testing.assert_array_equal(res.variables.cpu().numpy(), x)
    np.testing.assert_array_equal(res.objectives.cpu().numpy(), f.T)
    np.testing.assert_array_equal(res.auxiliary_indices.cpu().numpy(), fr)
    assert res.active_count == 92 and res.total_ms > 0


def test_nsga3_run_larger_matches_oracle():
    alg, prob = _config0(5, seed=1, n=512, d=64, m=5)
    res = alg.run()
    bounds = np.tile(np.array([0.0, 1.0], np.float32), 64)
    x, f, cv, fr, _ = O.nsga3_run(512, 5, 64, 5, 1, bounds, prob.centers.numpy())
    np.testing.assert_array_equal(res.variables.cpu().numpy(), x)
    np.testing.assert_array_equal(res.auxiliary_indices.cpu().numpy(), fr)


def test_step_and_reset_and_views():
    alg, _ = _config0(10)
    alg.step()
    assert alg.generation == 1 and alg.initialized and not alg.finished
    pop = alg.population
    assert tuple(pop.variables.shape) == (92, 12) and tuple(pop.objectives.shape) == (92, 3)
    while not alg.finished:
        alg.step()
    r1 = alg.result()
    alg.reset()
    assert alg.generation == 0
    r2 = alg.run()
    assert tuple(r2.variables.shape) == (92, 12)
    assert not torch.equal(r1.variables, r2.variables)   # reset draws a new population


def test_python_problem_and_rvea():
    class Sphere(cm.PythonProblem):
        def __init__(self):
            super().__init__(dimension=8, objectives=2)

        def evaluate(self, variables, context):
            return torch.stack([(variables ** 2).sum(1), ((variables - 1) ** 2).sum(1)], 1)

    for algo in (cm.NSGA3, cm.RVEA):
        alg = algo(population_size=40, max_generations=15, problem=Sphere(), print_progress=False)
        res = alg.run()
        x = res.variables[: res.active_count]
        expect = torch.stack([(x ** 2).sum(1), ((x - 1) ** 2).sum(1)], 1)
        torch.testing.assert_close(res.objectives[: res.active_count], expect, rtol=2e-5, atol=2e-4)


def test_rvea_run_shapes_and_front_zero():
    alg = cm.RVEA(population_size=105, max_generations=30, problem=cm.QuadraticCenters(10, 3, seed=2),
                  print_progress=False)
    res = alg.run()
    assert 1 <= res.active_count <= 105
    fz = res.auxiliary_indices.cpu().numpy()
    assert len(fz) >= 1 and fz.max() < res.active_count
    f = res.objectives[: res.active_count].cpu().numpy().T
    fr, _ = O.nds(np.ascontiguousarray(f), np.zeros(res.active_count, np.float32))
    np.testing.assert_array_equal(np.sort(fz), np.where(fr == 0)[0])


def test_invalid_configurations():
    with pytest.raises(ValueError):
        cm.NSGA3(population_size=91, problem=cm.QuadraticCenters(4, 3), print_progress=False).initialize()
    with pytest.raises(ValueError):
        cm.NSGA3(population_size=20, problem=cm.QuadraticCenters(4, 3), print_progress=False,
                 initial_population=np.full((20, 4), 2.0, np.float32)).initialize()
