"""Python strategy objects (PythonMating, PythonCrossover, PythonMutation,
PythonReferenceDirections; reference adapters.cpp:196-338) run as hooks of
the native generation loop.  A Python strategy that restates a built-in one
must give the built-in run bit for bit: that pins the hook order, the
generation and seed each hook sees, the given-index SBX pass and the
mutation-only pass against the fused kernel."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402
from cuda_moea import _lib  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
SEED = 5


def _run(gens=10, n=92, **kw):
    kind = kw.pop("kind", "NSGA3")
    alg = cm.Algorithm(kind, population_size=n, max_generations=gens, problem=cm.QuadraticCenters(12, 3, seed=1),
                       seed=SEED, print_progress=False, **kw)
    return alg, alg.run()


def _same(r1, r2):
    np.testing.assert_array_equal(r1.variables.cpu().numpy(), r2.variables.cpu().numpy())
    np.testing.assert_array_equal(r1.objectives.cpu().numpy(), r2.objectives.cpu().numpy())
    np.testing.assert_array_equal(r1.auxiliary_indices.cpu().numpy(), r2.auxiliary_indices.cpu().numpy())


class NativeTournament(cm.PythonMating):
    """The built-in tournament through the C ABI: the winners the fused kernel draws."""

    def __init__(self):
        self.calls = []

    def initialize(self, info):
        self.info = info

    def select(self, parents, state, context):
        self.calls.append((context["generation"], state["active_count"]))
        _, idx = K.reproduce(parents["variables"], parents["bounds"].reshape(-1).contiguous(), rank=state["rank"],
                             cv=parents["constraints"], seed=SEED, generation=context["generation"])
        return idx


class CopyCrossover(cm.PythonCrossover):
    def apply(self, parents, parent_indices, context):
        return parents["variables"][parent_indices.long()]


class KeepMutation(cm.PythonMutation):
    def __init__(self, give_none):
        self.give_none = give_none

    def mutate(self, offspring, context):
        return None if self.give_none else offspring["variables"].clone()


def test_python_mating_equals_builtin_tournament():
    print("strategies: mating", flush=True)
    _, ref = _run()
    mating = NativeTournament()
    _, got = _run(mating=mating)
    _same(ref, got)
    assert [c[0] for c in mating.calls] == list(range(10)) and all(c[1] == 92 for c in mating.calls)
    assert mating.info["name"] == "NSGA-III" and mating.info["population_size"] == 92


def test_python_crossover_then_builtin_mutation_equals_sbx_probability_zero():
    print("strategies: crossover", flush=True)
    _, ref = _run(crossover=cm.SBX(probability=0.0))
    _, got = _run(crossover=CopyCrossover())
    _same(ref, got)


@pytest.mark.parametrize("give_none", [False, True])
def test_python_mutation_equals_no_mutation(give_none):
    print(f"strategies: mutation none={give_none}", flush=True)
    _, ref = _run(mutation=cm.NoMutation())
    _, got = _run(mutation=KeepMutation(give_none))
    _same(ref, got)


def test_all_three_python_strategies_together():
    print("strategies: all three", flush=True)
    _, ref = _run(crossover=cm.SBX(probability=0.0), mutation=cm.NoMutation())
    _, got = _run(mating=NativeTournament(), crossover=CopyCrossover(), mutation=KeepMutation(False))
    _same(ref, got)


class FixedDasDennis(cm.PythonReferenceDirections):
    def initialize(self, requested_count, objective_count, context):
        w = torch.from_numpy(K.das_dennis(requested_count, objective_count))
        return w.t().contiguous().to(f"cuda:{context['device']}")


class PythonRveaAdaptation(cm.PythonReferenceDirections):
    """AdaptiveRVEADirections restated with the C ABI's adaptation kernel."""

    def __init__(self, frequency, max_generations):
        self.interval = max(1, round(frequency * max_generations))
        self.resets = 0

    def initialize(self, requested_count, objective_count, context):
        self.v0 = torch.from_numpy(K.das_dennis(requested_count, objective_count, 0, kind=2)).to(
            f"cuda:{context['device']}")
        return self.v0.t()

    def update(self, population, directions, active, context):
        assert directions.shape == (self.v0.shape[1], self.v0.shape[0])
        if (context["generation"] + 1) % self.interval != 0:
            return None
        f = population["objectives"].t()                     # the (m, n) storage
        return K.rvea_adapt(self.v0, f, active).t()

    def reset(self):
        self.resets += 1


def test_python_directions_equal_builtin_das_dennis():
    print("strategies: fixed directions", flush=True)
    _, ref = _run()
    _, got = _run(reference_directions=FixedDasDennis())
    _same(ref, got)


def test_python_directions_update_equals_rvea_adaptation():
    print("strategies: rvea adaptation", flush=True)
    a1, ref = _run(gens=12, n=100, kind="RVEA", reference_directions=cm.AdaptiveRVEADirections(frequency=0.25))
    prov = PythonRveaAdaptation(0.25, 12)
    a2, got = _run(gens=12, n=100, kind="RVEA", reference_directions=prov)
    _same(ref, got)
    assert ref.active_count == got.active_count
    np.testing.assert_array_equal(a1.directions.cpu().numpy(), a2.directions.cpu().numpy())
    a2.reset()
    assert prov.resets == 1
    np.testing.assert_array_equal(a2.directions.cpu().numpy(), prov.v0.cpu().numpy())


def test_reference_custom_strategy_classes_run():
    """The custom strategies of the reference's tests/python/test_api.py:17-36
    (identity mating, gathered crossover, pass-through mutation, identity
    directions) on a two-objective Python problem."""
    class Sphere(cm.PythonProblem):
        def __init__(self):
            super().__init__(4, 2)

        def evaluate(self, variables, context):
            return torch.stack((variables.square().sum(1), (variables - 1).square().sum(1)), 1)

    class Mating(cm.PythonMating):
        def select(self, parents, state, context):
            return torch.arange(parents["variables"].shape[0], device=parents["variables"].device,
                                dtype=torch.int32)

    class References(cm.PythonReferenceDirections):
        def initialize(self, requested_count, objective_count, context):
            return torch.eye(objective_count, device=f"cuda:{context['device']}")

    print("strategies: reference classes", flush=True)
    alg = cm.NSGA3(population_size=32, max_generations=3, problem=Sphere(), mating=Mating(),
                   crossover=CopyCrossover(), mutation=KeepMutation(False), reference_directions=References(),
                   print_progress=False, seed=3)
    x0 = None
    alg.initialize()
    x0 = alg.population.variables.clone()
    res = alg.run()
    assert tuple(res.variables.shape) == (32, 4) and tuple(res.objectives.shape) == (32, 2)
    assert alg.directions.shape == (2, 2)
    # identity pairs copied, no mutation: every survivor is an initial row
    rows = {tuple(r) for r in x0.cpu().numpy().tolist()}
    assert all(tuple(r) in rows for r in res.variables.cpu().numpy().tolist())


class BadMating(cm.PythonMating):
    def __init__(self, kind):
        self.kind = kind

    def select(self, parents, state, context):
        n = parents["variables"].shape[0]
        dev = parents["variables"].device
        if self.kind == "count":
            return torch.zeros(n - 1, dtype=torch.int32, device=dev)
        if self.kind == "range":
            return torch.full((n,), n, dtype=torch.int32, device=dev)
        return np.zeros(n, np.int32)


@pytest.mark.parametrize("kind,match", [("count", "wrong number"), ("range", r"\[0, population_size\)"),
                                        ("host", "torch.Tensor")])
def test_bad_mating_output_raises(kind, match):
    alg = cm.NSGA3(population_size=92, max_generations=3, problem=cm.QuadraticCenters(12, 3), mating=BadMating(kind),
                   print_progress=False)
    with pytest.raises(ValueError, match=match):
        alg.step()


def test_bad_crossover_and_direction_outputs_raise():
    class Short(cm.PythonCrossover):
        def apply(self, parents, parent_indices, context):
            return parents["variables"][:-2]

    class Grow(cm.PythonReferenceDirections):
        def initialize(self, requested_count, objective_count, context):
            return torch.eye(objective_count, device=f"cuda:{context['device']}")

        def update(self, population, directions, active, context):
            return torch.cat([directions, directions])

    alg = cm.NSGA3(population_size=92, max_generations=3, problem=cm.QuadraticCenters(12, 3), crossover=Short(),
                   print_progress=False)
    with pytest.raises(ValueError, match="crossover output"):
        alg.step()
    alg = cm.NSGA3(population_size=92, max_generations=3, problem=cm.QuadraticCenters(12, 3),
                   reference_directions=Grow(), print_progress=False)
    with pytest.raises(ValueError, match="cannot change count"):
        alg.step()


def test_reproduce_given_indices_match_computed_winners():
    """MOEA_MATING_GIVEN through moea_reproduce: the winners the tournament
    computed, handed back, give the same offspring."""
    g = torch.Generator().manual_seed(0)
    x = torch.rand(300, 7, generator=g).cuda()
    bounds = torch.tensor([0.0, 1.0] * 7).cuda()
    rank = torch.randint(0, 4, (300,), generator=g, dtype=torch.int32).cuda()
    cv = torch.zeros(300).cuda()
    off, idx = K.reproduce(x, bounds, rank=rank, cv=cv, seed=9, generation=4)
    off2 = torch.empty_like(off)
    a = _lib.ReproduceArgs()
    a.n, a.d, a.active, a.mating = 300, 7, 300, _lib.MATING_GIVEN
    a.parents_x, a.bounds = x.data_ptr(), bounds.data_ptr()
    a.sbx_eta, a.sbx_prob, a.sbx_copy_prob = 30.0, 1.0, 0.0
    a.mutation, a.pm_eta, a.pm_prob = _lib.MUTATION_POLYNOMIAL, 20.0, 1.0
    a.seed, a.generation = 9, 4
    a.offspring_x, a.parent_indices = off2.data_ptr(), idx.data_ptr()
    a.stream = torch.cuda.current_stream().cuda_stream
    assert _lib.lib.moea_reproduce(ctypes.byref(a)) == 0
    torch.cuda.synchronize()
    assert torch.equal(off, off2)


class PythonNsga3Selector(cm.PythonEnvironmentSelector):
    """The built-in NSGA-III selection through the C ABI, returned as arrays
    with the survivors' ranks as the mating state and the result indices."""

    def initialize(self, info, references, context):
        self.info, self.references = info, references

    def prepare(self, population, context):
        f = population["objectives"].t().contiguous()
        self.ideal = K.ideal_point(f)
        rank, _ = K.nds_rank(f, population["constraints"].contiguous())
        return {"rank": rank, "result_indices": rank}

    def select(self, parents, offspring, context):
        def pop(d):
            return K.Population(d["variables"], d["objectives"].t().contiguous(), d["constraints"])
        w = self.references.t().contiguous()
        nxt, nfr, _ = K.nsga3_select(pop(parents), pop(offspring), w, self.ideal, seed=SEED,
                                     generation=context["generation"])
        return {"variables": nxt.x, "objectives": nxt.f.t(), "constraints": nxt.cv, "rank": nfr,
                "result_indices": nfr}


def test_python_selector_equals_builtin_nsga3_selection():
    print("strategies: nsga3 selector", flush=True)
    _, ref = _run()
    sel = PythonNsga3Selector()
    _, got = _run(environment_selector=sel)
    _same(ref, got)
    assert got.active_count == 92 and sel.info["name"] == "NSGA-III" and sel.references.shape == (91, 3)


class PythonRveaSelector(cm.PythonEnvironmentSelector):
    """The built-in RVEA selection through the C ABI; the engine still adapts
    the directions, with the active count this selector reports."""

    def initialize(self, info, references, context):
        self.references, self.max_gen = references, info["max_generations"]
        self.active = info["population_size"]

    def prepare(self, population, context):
        self.active = population["variables"].shape[0]
        return {"active_count": self.active}

    def select(self, parents, offspring, context):
        def pop(d):
            return K.Population(d["variables"], d["objectives"].t().contiguous(), d["constraints"])
        v = self.references.t().contiguous()
        m = v.shape[0]
        pen = K.rvea_penalty(m, context["generation"] + 1, self.max_gen, 2.0)
        nxt, self.active = K.rvea_select(pop(parents), pop(offspring), v, K.rvea_gamma(v), self.active, pen)
        return {"variables": nxt.x, "objectives": nxt.f.t(), "constraints": nxt.cv, "active_count": self.active}


def test_python_selector_equals_builtin_rvea_selection():
    print("strategies: rvea selector", flush=True)
    a1, ref = _run(gens=12, n=100, kind="RVEA", reference_directions=cm.AdaptiveRVEADirections(frequency=0.25))
    a2, got = _run(gens=12, n=100, kind="RVEA", reference_directions=cm.AdaptiveRVEADirections(frequency=0.25),
                   environment_selector=PythonRveaSelector())
    np.testing.assert_array_equal(ref.variables.cpu().numpy(), got.variables.cpu().numpy())
    np.testing.assert_array_equal(ref.objectives.cpu().numpy(), got.objectives.cpu().numpy())
    assert ref.active_count == got.active_count
    np.testing.assert_array_equal(a1.directions.cpu().numpy(), a2.directions.cpu().numpy())


class IndexSelector(cm.PythonEnvironmentSelector):
    """The reference test's selector (tests/python/test_api.py:39-43) and its
    offspring-only twin: survivors by index into parents ++ offspring."""

    def __init__(self, offspring_only):
        self.offspring_only = offspring_only
        self.resets = 0
        self.seen = None

    def select(self, parents, offspring, context):
        n = parents["variables"].shape[0]
        self.seen = offspring["variables"].clone()
        lo = n if self.offspring_only else 0
        return {"indices": torch.arange(lo, lo + n, device=parents["variables"].device),
                "score": torch.ones(n, device=parents["variables"].device), "active_count": n - 2}

    def reset(self):
        self.resets += 1


class StateMating(cm.PythonMating):
    def select(self, parents, state, context):
        self.keys = sorted(state)
        self.active = state["active_count"]
        n = parents["variables"].shape[0]
        return torch.arange(n, dtype=torch.int32, device=parents["variables"].device)


@pytest.mark.parametrize("offspring_only", [False, True])
def test_python_selector_indices_and_mating_state(offspring_only):
    print(f"strategies: index selector offspring_only={offspring_only}", flush=True)
    sel, mating = IndexSelector(offspring_only), StateMating()
    alg = cm.NSGA3(population_size=92, max_generations=4, problem=cm.QuadraticCenters(12, 3, seed=1),
                   mating=mating, environment_selector=sel, print_progress=False, seed=2)
    alg.initialize()
    x0 = alg.population.variables.clone()
    alg.step()
    want = sel.seen if offspring_only else x0
    assert torch.equal(alg.population.variables, want)
    alg.step()
    assert mating.keys == ["active_count", "score"] and mating.active == 90
    res = alg.run()
    assert res.active_count == 90 and res.auxiliary_indices.numel() == 0
    alg.reset()
    assert sel.resets == 1 and alg.generation == 0


def test_tournament_needs_the_selector_rank():
    alg = cm.NSGA3(population_size=92, max_generations=3, problem=cm.QuadraticCenters(12, 3),
                   environment_selector=IndexSelector(False), print_progress=False)
    with pytest.raises(ValueError, match="rank array"):
        alg.initialize()


def test_selector_bad_indices_raise():
    class Far(cm.PythonEnvironmentSelector):
        def select(self, parents, offspring, context):
            n = parents["variables"].shape[0]
            return {"indices": torch.full((n,), 2 * n, device=parents["variables"].device)}

    alg = cm.NSGA3(population_size=92, max_generations=3, problem=cm.QuadraticCenters(12, 3), mating=cm.RandomMating(),
                   environment_selector=Far(), print_progress=False)
    with pytest.raises(ValueError, match=r"\[0, 2 \* population_size\)"):
        alg.step()
