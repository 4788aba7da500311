"""Whole RVEA runs: the native loop (random mating over the live parents, SBX,
PM, quadratic-center evaluation, RVEA selection with t = generation + 1,
dead-row sentinels, direction adaptation every round(fr x G) generations)
against the oracle's run restatement, bit for bit; and the full-size RVEA
configuration (N = 16384, M = 5, D = 131072, seeded MLP D -> 256 -> 5) checked
through properties that hold at any size."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
import cuda_moea as cm  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _rvea(n, m, d, gens, seed, freq, alpha):
    prob = cm.QuadraticCenters(d, m, seed=seed)
    alg = cm.RVEA(population_size=n, max_generations=gens, problem=prob, seed=seed,
                  environment_selector=cm.RVEAEnvironmentSelector(alpha=alpha),
                  reference_directions=cm.AdaptiveRVEADirections(frequency=freq), print_progress=False)
    return alg, prob


@pytest.mark.parametrize("n,m,d,gens,seed,freq,alpha", [
    (120, 5, 12, 20, 3, 0.1, 2.0),       # adaptation after generations 1, 3, 5, ... (interval 2)
    (92, 3, 12, 30, 0, 0.2, 2.0),        # interval 6
    (210, 8, 40, 12, 7, 0.25, 1.5),      # interval 3, another alpha
    (64, 2, 8, 25, 11, 1.0, 2.0),        # one adaptation, after the last generation
])
def test_rvea_run_matches_oracle(n, m, d, gens, seed, freq, alpha):
    alg, prob = _rvea(n, m, d, gens, seed, freq, alpha)
    print(f"rvea run n={n} m={m} d={d} gens={gens} seed={seed} fr={freq} alpha={alpha}", flush=True)
    alg.initialize()
    v0 = alg._directions0.cpu().numpy()
    res = alg.run()
    bounds = np.tile(np.array([0.0, 1.0], np.float32), d)
    x, f, cv, v, active = O.rvea_run(n, m, d, v0, gens, seed, bounds, prob.centers.numpy(), alpha=alpha,
                                     frequency=freq)
    assert res.active_count == active
    np.testing.assert_array_equal(res.variables.cpu().numpy(), x)       # NaN rows compare equal
    np.testing.assert_array_equal(res.objectives.cpu().numpy(), f.T)
    np.testing.assert_array_equal(res.constraints.cpu().numpy(), cv)
    np.testing.assert_array_equal(alg._directions.cpu().numpy(), v)
    # the dead rows hold the sentinels
    assert np.all(np.isnan(x[active:])) and np.all(f[:, active:] == 1e15) and np.all(cv[active:] == 1e15)


def test_rvea_run_step_by_step_matches_oracle_prefixes():
    """Generation by generation: the state after every step equals the oracle
    run cut after that many generations (the adaptation schedule, the live
    count and the sentinels in between)."""
    n, m, d, gens, seed = 60, 3, 10, 8, 5
    alg, prob = _rvea(n, m, d, gens, seed, 0.25, 2.0)
    alg.initialize()
    v0 = alg._directions0.cpu().numpy()
    bounds = np.tile(np.array([0.0, 1.0], np.float32), d)
    for g in range(1, gens + 1):
        alg.step()
        print(f"rvea step {g}", flush=True)
        pop = alg.population
        x, f, cv, v, active = O.rvea_run(n, m, d, v0, gens, seed, bounds, prob.centers.numpy(),
                                         frequency=0.25, limit=g)
        np.testing.assert_array_equal(pop.variables.cpu().numpy(), x)
        np.testing.assert_array_equal(pop.objectives.cpu().numpy(), f.T)
        np.testing.assert_array_equal(alg._directions.cpu().numpy(), v)
        assert int(alg._active.item()) == active


def test_rvea_config2_full_size_properties():
    """The RVEA configuration at its full size: N = 16384, M = 5, D = 131072,
    seeded MLP D -> 256 -> 5, adaptive directions (fr 0.1), alpha 2, three
    generations of the 200-generation schedule."""
    n, m, d = 16384, 5, 131072
    prob = cm.SeededMLP(d, m, hidden=256, seed=0, device=DEV)
    alg = cm.RVEA(population_size=n, max_generations=200, problem=prob, seed=0,
                  environment_selector=cm.RVEAEnvironmentSelector(alpha=2.0),
                  reference_directions=cm.AdaptiveRVEADirections(frequency=0.1), print_progress=False)
    alg.initialize()
    k = alg.k
    for g in range(3):
        print(f"rvea config2 generation {g}", flush=True)
        alg.step()
    res = alg.result()
    active = res.active_count
    x, f, cv = res.variables, res.objectives, res.constraints
    assert 0 < active <= min(k, n)
    # live rows finite and in bounds, dead rows sentinels
    assert torch.isfinite(f[:active]).all() and torch.isfinite(x[:active]).all()
    assert bool(((x[:active] >= 0) & (x[:active] <= 1)).all())
    assert bool(torch.isnan(x[active:]).all()) and bool((f[active:] == 1e15).all())
    assert bool((cv[active:] == 1e15).all()) and bool((cv[:active] == 0).all())
    # objectives of a sample of survivors equal the MLP of their variables in
    # float64: W1 x sums d = 131072 float32 products of size ~2^-8.5, so its
    # absolute error is ~ sqrt(d) 2^-24 per hidden unit (stated bound 64 x that,
    # carried through tanh and W2 into y and y^2, relative to 1 + |f|)
    rows = torch.linspace(0, active - 1, min(active, 48)).round().long().unique().to(DEV)
    xs = x[rows].double()
    h = torch.tanh(xs @ prob.w1.double().t())
    y = h @ prob.w2.double().t()
    ref = y ** 2 + 0.01 * ((xs - 0.5) ** 2).sum(1, keepdim=True)
    got = f[rows].double()
    tol = 64 * np.sqrt(d) * 2.0 ** -24 * (1.0 + ref.abs())
    assert bool(((got - ref).abs() <= tol).all()), float(((got - ref).abs() / (1 + ref.abs())).max())
