"""Snapshot recording of a run (save_data / save_interval) against the
oracle's trajectory: each saved generation holds that generation's
population, bit for bit, in the reference's on-disk layout
(python/csrc/native/src/algorithm/data_recorder.cu:129-290, call sites
algorithm.cu:208-214, 271-273, 336-339, 430)."""
import json

import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
import cuda_moea as cm  # noqa: E402
from cuda_moea.record import load_snapshot  # noqa: E402

pytestmark = pytest.mark.gpu
BOUNDS = np.tile(np.array([0.0, 1.0], np.float32), 12)


def _saved(root):
    return sorted(p.name for p in root.iterdir() if p.name.startswith("generation_"))


def _check_generation(root, g, prob, gens, seed):
    snap = load_snapshot(root, g)
    if g == 0:
        x, f, cv, fr, _ = O.nsga3_run(92, 3, 12, 0, seed, BOUNDS, prob.centers.numpy())
    else:
        x, f, cv, fr, _ = O.nsga3_run(92, 3, 12, gens, seed, BOUNDS, prob.centers.numpy(), limit=g)
    np.testing.assert_array_equal(snap["variables"], x, err_msg=f"generation {g}")
    np.testing.assert_array_equal(snap["objectives"], f)
    np.testing.assert_array_equal(snap["constraints"], cv)
    np.testing.assert_array_equal(snap["auxiliary_indices"], fr)
    assert snap["generation"] == g and snap["active_count"] == 92 and snap["auxiliary_index_count"] == 92
    np.testing.assert_array_equal(snap["reference_directions"], O.das_dennis(92, 3))


@pytest.mark.parametrize("driver", ["run", "step"])
def test_nsga3_snapshots_follow_the_oracle_trajectory(tmp_path, driver):
    gens, seed = 7, 4
    prob = cm.QuadraticCenters(12, 3, seed=seed)
    root = tmp_path / "rec"
    alg = cm.NSGA3(population_size=92, max_generations=gens, problem=prob, seed=seed,
                   print_progress=False, save_data=root, save_interval=3)
    print(f"snapshot nsga3 driver={driver}", flush=True)
    if driver == "run":
        res = alg.run()
    else:
        while not alg.finished:
            alg.step()
        res = alg.result()
    assert _saved(root) == ["generation_000000", "generation_000003", "generation_000006", "generation_000007"]
    meta = json.loads((root / "metadata.json").read_text())
    assert meta["algorithm_name"] == "NSGA-III" and meta["problem_name"] == "QuadraticCenters"
    assert (meta["population_size"], meta["dimension"], meta["objective_count"]) == (92, 12, 3)
    assert meta["max_generations"] == gens and meta["random_seed"] == seed and meta["generation_interval"] == 3
    assert meta["reference_direction_count"] == 91 and meta["constraint_count"] == 0
    np.testing.assert_array_equal(np.fromfile(root / "lower_bounds.bin", "<f4"), np.zeros(12, np.float32))
    np.testing.assert_array_equal(np.fromfile(root / "upper_bounds.bin", "<f4"), np.ones(12, np.float32))
    np.testing.assert_array_equal(np.fromfile(root / "initial_reference_directions.bin", "<f4").reshape(3, 91),
                                  O.das_dennis(92, 3))
    for g in (0, 3, 6, 7):
        _check_generation(root, g, prob, gens, seed)
    last = load_snapshot(root, gens)
    np.testing.assert_array_equal(last["variables"], res.variables.cpu().numpy())
    np.testing.assert_array_equal(last["objectives"], res.objectives.cpu().numpy().T)
    # reset starts a new trajectory and rewrites generation 0
    alg.reset()
    snap0 = load_snapshot(root, 0)
    np.testing.assert_array_equal(snap0["variables"], alg.population.variables.cpu().numpy())


def test_rvea_snapshots_directions_and_front_zero(tmp_path):
    root = tmp_path / "rvea"
    prob = cm.QuadraticCenters(12, 3, seed=2)
    alg = cm.RVEA(population_size=100, max_generations=5, problem=prob, seed=2, print_progress=False,
                  save_data=str(root), save_interval=2,
                  reference_directions=cm.AdaptiveRVEADirections(frequency=0.4))
    print("snapshot rvea", flush=True)
    alg.step()
    alg.step()
    g2 = load_snapshot(root, 2)
    # no result() yet: the selector holds no front-zero indices (rvea_environment_selector.cu:299-305)
    assert g2["auxiliary_indices"] is None and not (root / "generation_000002" /
                                                     "selector_auxiliary_indices.bin").exists()
    np.testing.assert_array_equal(g2["reference_directions"], alg.directions.cpu().numpy())
    np.testing.assert_array_equal(g2["variables"], alg.population.variables.cpu().numpy())
    assert g2["active_count"] == int(alg._active.item())
    res = alg.run()
    g5 = load_snapshot(root, 5)
    assert _saved(root) == ["generation_000000", "generation_000002", "generation_000004", "generation_000005"]
    np.testing.assert_array_equal(g5["auxiliary_indices"], res.auxiliary_indices.cpu().numpy())
    assert g5["active_count"] == res.active_count
    np.testing.assert_array_equal(g5["reference_directions"], alg.directions.cpu().numpy())
    init = np.fromfile(root / "initial_reference_directions.bin", "<f4").reshape(3, -1)
    np.testing.assert_array_equal(init, load_snapshot(root, 0)["reference_directions"])
    assert json.loads((root / "metadata.json").read_text())["algorithm_name"] == "RVEA"


def test_snapshot_interval_must_be_positive(tmp_path):
    with pytest.raises(ValueError, match="interval"):
        cm.NSGA3(population_size=92, max_generations=3, save_data=tmp_path, save_interval=0)
