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


def test_python_problem_prepare_parent_and_reset():
    """IProblemEvaluator::prepare_parent runs before every reproduction
    (algorithm.cu:230-241): an evaluation it returns replaces the parents'
    objectives and the NSGA-III parent fronts are re-sorted; True reports an
    in-place change; reset() reaches the problem (algorithm.cu:413)."""
    class Dynamic(cm.PythonProblem):
        def __init__(self, in_place):
            super().__init__(dimension=6, objectives=2)
            self.calls, self.resets, self.in_place = [], 0, in_place

        def evaluate(self, x, context):
            return torch.stack([(x ** 2).sum(1), ((x - 1) ** 2).sum(1)], 1)

        def prepare_parent(self, population, context):
            self.calls.append(context["generation"])
            f = self.evaluate(population["variables"], context)
            f[0] = -1e6                     # parent 0 now dominates every other row
            if self.in_place:
                population["objectives"].copy_(f)
                return True
            return {"objectives": f, "constraints": torch.zeros_like(f[:, 0])}

        def reset(self):
            self.resets += 1

    for in_place in (False, True):
        prob = Dynamic(in_place)
        alg = cm.NSGA3(population_size=40, max_generations=5, problem=prob, seed=2, print_progress=False)
        print(f"prepare_parent in_place={in_place}", flush=True)
        for g in range(5):
            alg.step()
            # merged row 0 (parent 0) is the only first-front row: it survives
            # as row 0 of the next population with its rewritten objectives, and
            # the re-sorted parent fronts ranked it alone in front 0
            f = alg.population.objectives.cpu().numpy()
            assert np.all(f[0] == -1e6) and np.all(f[1:] > -1e6)
        assert prob.calls == [0, 1, 2, 3, 4]
        res = alg.result()
        assert int((res.auxiliary_indices == 0).sum()) == 1
        alg.reset()
        assert prob.resets == 1
