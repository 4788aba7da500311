"""Batched NSGA-III (moea_batch_*): every run of a batch equals the same run
alone, bit for bit, and the oracle's trajectory for the config-0 shape."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
import cuda_moea as cm  # noqa: E402

pytestmark = pytest.mark.gpu

# (n, m, d, partitions, seed): uneven populations, objective counts and dimensions
MIX = [(92, 3, 12, 12, 0), (64, 5, 30, 6, 1), (130, 8, 7, 3, 2), (256, 2, 100, 20, 3),
       (200, 3, 64, 0, 4), (1000, 4, 50, 0, 5), (2, 3, 5, 12, 6)]


# seeded-MLP instances: (n, m, d, partitions, seed, hidden); hidden need not be a multiple of 64
MIX_MLP = [(96, 3, 40, 12, 7, 256), (64, 5, 130, 6, 8, 100), (130, 2, 9, 20, 9, 64)]


def _run(n, m, d, parts, seed, gens):
    return cm.NSGA3(population_size=n, max_generations=gens,
                    problem=cm.QuadraticCenters(d, m, seed=seed + 100), seed=seed,
                    reference_directions=cm.DasDennisDirections(partitions=parts),
                    print_progress=False)


def _run_mlp(n, m, d, parts, seed, hidden, gens):
    return cm.NSGA3(population_size=n, max_generations=gens,
                    problem=cm.SeededMLP(d, m, hidden=hidden, seed=seed + 200), seed=seed,
                    reference_directions=cm.DasDennisDirections(partitions=parts),
                    print_progress=False)


def _same(a, b):
    np.testing.assert_array_equal(a.variables.cpu().numpy(), b.variables.cpu().numpy())
    np.testing.assert_array_equal(a.objectives.cpu().numpy(), b.objectives.cpu().numpy())
    np.testing.assert_array_equal(a.constraints.cpu().numpy(), b.constraints.cpu().numpy())
    np.testing.assert_array_equal(a.auxiliary_indices.cpu().numpy(), b.auxiliary_indices.cpu().numpy())


@pytest.mark.parametrize("gens", [1, 15])
def test_batch_runs_equal_single_runs(gens):
    print(f"batch mix gens={gens}", flush=True)
    batch = cm.BatchNSGA3([_run(*c, gens) for c in MIX])
    results = batch.run()
    for c, res in zip(MIX, results):
        print(f"  instance {c}", flush=True)
        _same(res, _run(*c, gens).run())


def test_batch_mixed_objectives_equal_single_runs():
    gens = 8
    runs = [_run(*c, gens) for c in MIX[:4]] + [_run_mlp(*c, gens) for c in MIX_MLP]
    results = cm.BatchNSGA3(runs).run()
    singles = [_run(*c, gens) for c in MIX[:4]] + [_run_mlp(*c, gens) for c in MIX_MLP]
    for i, (res, single) in enumerate(zip(results, singles)):
        print(f"  instance {i}", flush=True)
        _same(res, single.run())


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_niching_paths_equal_single_runs(mode):
    """The batch's niching paths (one workgroup per instance, chosen while every
    2n <= 8192; device-wide sorts otherwise or in mode 2), and its pruned
    association for instances with K >= 2048, against single runs."""
    gens = 6
    # 2n = 8400: device-wide niching in mode 0; K >= 2048: pruned association
    cases = MIX[:3] + [(4200, 3, 20, 0, 11), (2600, 5, 15, 0, 12)]
    print(f"batch niching mode {mode}", flush=True)
    results = cm.BatchNSGA3([_run(*c, gens) for c in cases], niching=(mode, 0)).run()
    for c, res in zip(cases, results):
        print(f"  instance {c}", flush=True)
        _same(res, _run(*c, gens).run())


def test_batch_alone_equals_in_mix():
    gens = 10
    alone = cm.BatchNSGA3([_run(*MIX[5], gens)]).run()[0]
    mixed = cm.BatchNSGA3([_run(*c, gens) for c in MIX]).run()[5]
    _same(alone, mixed)


def test_batch_config0_matches_oracle():
    gens = 30
    runs = [_run(*c, gens) for c in MIX]
    res = cm.BatchNSGA3(runs).run()[0]
    prob = runs[0].problem
    bounds = np.tile(np.array([0.0, 1.0], np.float32), 12)
    x, f, cv, fr, _ = O.nsga3_run(92, 3, 12, gens, 0, bounds, prob.centers.numpy())
    np.testing.assert_array_equal(res.variables.cpu().numpy(), x)
    np.testing.assert_array_equal(res.objectives.cpu().numpy(), f.T)
    np.testing.assert_array_equal(res.auxiliary_indices.cpu().numpy(), fr)


def test_batch_step_and_validation():
    b = cm.BatchNSGA3([_run(*c, 5) for c in MIX[:3]])
    b.step(2)
    assert b.generation == 2 and not b.finished
    b.step(10)
    assert b.generation == 5 and b.finished
    with pytest.raises(ValueError):
        cm.BatchNSGA3([_run(92, 3, 12, 12, 0, 5), _run(92, 3, 12, 12, 0, 6)])
    with pytest.raises(ValueError):
        cm.BatchNSGA3([_run(91, 3, 12, 12, 0, 5)])
