"""Native DTLZ1 / DTLZ2 objectives (the reference's built-in kernels,
problem/dtlz/dtlz_kernels.cu:28-190) against an FP64 restatement, at the
reference's native-objective tolerance (run_benchmark.py:54-61: rtol 2e-5,
atol 2e-4)."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402
from cuda_moea import _lib, kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def dtlz_fp64(kind, x, m):
    x = x.astype(np.float64)
    xm, xk = x[:, : m - 1], x[:, m - 1:]
    if kind == 1:
        g = 100.0 * (xk.shape[1] + ((xk - 0.5) ** 2 - np.cos(20 * math.pi * (xk - 0.5))).sum(1))
        base = 0.5 * (1 + g)
    else:
        base = 1.0 + ((xk - 0.5) ** 2).sum(1)
    f = np.empty((m, x.shape[0]))
    for i in range(m):
        v = base.copy()
        for j in range(m - 1 - i):
            v *= np.clip(xm[:, j], 0, 1) if kind == 1 else np.maximum(0.0, np.cos(xm[:, j] * math.pi / 2))
        if i > 0:
            xi = xm[:, m - 1 - i]
            v *= np.maximum(0.0, 1 - xi) if kind == 1 else np.sin(xi * math.pi / 2)
        f[i] = v
    return f


@pytest.mark.parametrize("kind,n,d,m", [(1, 300, 12, 3), (1, 1000, 500, 3), (1, 77, 40, 8),
                                        (2, 300, 12, 3), (2, 513, 300, 5), (2, 64, 10, 10)])
def test_dtlz_kernel_against_fp64(kind, n, d, m):
    rng = np.random.default_rng(kind * 100 + m)
    x = rng.random((n, d), dtype=np.float32)
    print(f"dtlz{kind} n={n} d={d} m={m}", flush=True)
    f, cv = K.eval_dtlz(kind, torch.from_numpy(x).to(DEV), m)
    np.testing.assert_allclose(f.cpu().numpy(), dtlz_fp64(kind, x, m), rtol=2e-5, atol=2e-4)
    assert not cv.cpu().numpy().any()


@pytest.mark.parametrize("cls,kind", [(cm.DTLZ1, 1), (cm.DTLZ2, 2)])
def test_dtlz_run_native(cls, kind):
    alg = cm.NSGA3(population_size=92, max_generations=20, problem=cls(dimension=12, objectives=3),
                   seed=3, print_progress=False)
    alg.initialize()
    assert alg._args.problem.kind == (_lib.PROBLEM_DTLZ1 if kind == 1 else _lib.PROBLEM_DTLZ2)
    res = alg.run()
    x = res.variables.cpu().numpy()
    np.testing.assert_allclose(res.objectives.cpu().numpy().T, dtlz_fp64(kind, x, 3), rtol=2e-5, atol=2e-4)


def dtlz_fp64_suite(kind, x, m):
    """FP64 DTLZ3-7 (kinds 3-7) and ConvexDTLZ2 (kind 8) as the reference
    defines them (dtlz_kernels.cu:194-717): DTLZ4 maps x to x^100, DTLZ5/6 map
    x_j (j >= 1) to (1 + 2 g x_j) / (2 (1 + g)), DTLZ5 clamping it to [0, 1]."""
    x = x.astype(np.float64)
    n = x.shape[0]
    xm, xk = x[:, : m - 1], x[:, m - 1:]
    k = xk.shape[1]
    f = np.empty((m, n))
    if kind == 7:
        g = 1.0 + 9.0 / k * xk.sum(1)
        f[: m - 1] = xm.T
        h = m - (xm / (1.0 + g)[:, None] * (1.0 + np.sin(3 * math.pi * xm))).sum(1)
        f[m - 1] = (1.0 + g) * h
        return f
    if kind == 3:
        g = 100.0 * (k + ((xk - 0.5) ** 2 - np.cos(20 * math.pi * (xk - 0.5))).sum(1))
    elif kind == 6:
        g = (xk ** 0.1).sum(1)
    else:
        g = ((xk - 0.5) ** 2).sum(1)
    th = xm.copy()
    if kind == 4:
        th = th ** 100
    if kind in (5, 6):
        t = (1.0 + 2.0 * g[:, None] * xm) / (2.0 * (1.0 + g[:, None]))
        th[:, 1:] = np.clip(t, 0.0, 1.0)[:, 1:] if kind == 5 else t[:, 1:]
    ang = th * math.pi / 2
    for i in range(m):
        v = 1.0 + g
        for j in range(m - 1 - i):
            v = v * np.maximum(0.0, np.cos(ang[:, j]))
        if i > 0:
            v = v * np.sin(ang[:, m - 1 - i])
        f[i] = v
    if kind == 8:
        f[: m - 1] = f[: m - 1] ** 4
        f[m - 1] = f[m - 1] ** 2
    return f


SUITE = {3: cm.DTLZ3, 4: cm.DTLZ4, 5: cm.DTLZ5, 6: cm.DTLZ6, 7: cm.DTLZ7, 8: cm.ConvexDTLZ2}


@pytest.mark.parametrize("kind", sorted(SUITE))
@pytest.mark.parametrize("n,d,m", [(300, 12, 3), (513, 300, 5), (64, 20, 10)])
def test_dtlz_suite_kernels_against_fp64(kind, n, d, m):
    rng = np.random.default_rng(kind * 1000 + m)
    x = rng.random((n, d), dtype=np.float32)
    print(f"dtlz kind={kind} n={n} d={d} m={m}", flush=True)
    f, cv = K.eval_dtlz(kind, torch.from_numpy(x).to(DEV), m)
    _close_scaled(f.cpu().numpy(), dtlz_fp64_suite(kind, x, m), m)
    assert not cv.cpu().numpy().any()


def _close_scaled(got, ref, m):
    # The reference's tolerance (rtol 2e-5, atol 2e-4) plus the float32
    # conditioning of the formula: each angle is rounded to float32, which moves
    # its cos / sin factor by up to about 1e-7, times the row's scale (1 + g, up
    # to 3e4 for DTLZ3 at d = 300), for each of the m - 1 factors.
    scale = np.abs(ref).max(axis=0, keepdims=True)
    tol = 2e-5 * np.abs(ref) + 2e-4 + 4e-7 * m * scale
    bad = np.abs(got - ref) > tol
    assert not bad.any(), (np.argwhere(bad)[:5], got[bad][:5], ref[bad][:5])


@pytest.mark.parametrize("kind", [2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("m", [3, 6])
def test_dtlz_pareto_front_identities(kind, m):
    """On the Pareto set (g at its minimum) the objectives satisfy the
    problem's front equation, independent of any restatement: sum f^2 = 1 for
    the spherical problems, f_M + sum sqrt(f_i) = 1 for ConvexDTLZ2, and
    f_M = 2 (M - sum f_i / 2 (1 + sin 3 pi f_i)) with g = 1 for DTLZ7."""
    rng = np.random.default_rng(7 + kind)
    n, d = 400, m + 9
    x = rng.random((n, d), dtype=np.float32)
    x[:, m - 1:] = 0.0 if kind in (6, 7) else 0.5
    f = K.eval_dtlz(kind, torch.from_numpy(x).to(DEV), m)[0].cpu().numpy().astype(np.float64)
    if kind == 8:
        np.testing.assert_allclose(f[m - 1] + np.sqrt(f[: m - 1]).sum(0), 1.0, atol=2e-5)
    elif kind == 7:
        xm = x[:, : m - 1].astype(np.float64)
        np.testing.assert_array_equal(f[: m - 1], xm.T)
        np.testing.assert_allclose(f[m - 1], 2.0 * (m - (xm / 2.0 * (1 + np.sin(3 * math.pi * xm))).sum(1)),
                                   rtol=1e-5, atol=1e-5)
    else:
        np.testing.assert_allclose((f ** 2).sum(0), 1.0, atol=2e-5)
        assert (f >= 0).all()


@pytest.mark.parametrize("kind", sorted(SUITE))
def test_dtlz_suite_runs_native(kind):
    alg = cm.NSGA3(population_size=92, max_generations=10, problem=SUITE[kind](dimension=12, objectives=3),
                   seed=3, print_progress=False)
    alg.initialize()
    assert alg._args.problem.kind == _lib.DTLZ_KINDS[SUITE[kind].__name__]
    res = alg.run()
    x = res.variables.cpu().numpy()
    _close_scaled(res.objectives.cpu().numpy().T, dtlz_fp64_suite(kind, x, 3), 3)


# constrained variant -> (moea_eval_dtlz kind, base kind); radii of dtlz_kernels.cu:1110-1158
CONSTRAINED = {"C1DTLZ1": (9, 1), "C1DTLZ3": (10, 3), "C2DTLZ2": (11, 2), "C2ConvexDTLZ2": (12, 8),
               "C3DTLZ1": (13, 1), "C3DTLZ4": (14, 4)}
RADIUS = {"C1DTLZ3": {2: 6.0, 3: 5.0, 5: 12.5, 8: 15.0, 10: 15.0},
          "C2DTLZ2": {2: 0.15, 3: 0.40, 5: 1.0, 8: 1.0, 10: 1.0},
          "C2ConvexDTLZ2": {3: 0.20, 5: 0.225, 8: 0.26, 10: 0.26}}


def cv_fp64(name, f):
    """The constraint violations of C1/C2/C3 (Jain and Deb 2014, as
    dtlz_kernels.cu:824-1044 writes them), in FP64 from objectives f (m, n)."""
    f = f.astype(np.float64)
    m = f.shape[0]
    r = np.float64(np.float32(RADIUS.get(name, {}).get(m, 1.0)))
    s2 = (f ** 2).sum(0)
    if name == "C1DTLZ1":
        return np.maximum(f[m - 1] / 0.6 + (f[: m - 1] / 0.5).sum(0) - 1.0, 0.0)
    if name == "C1DTLZ3":
        return np.maximum((s2 - 16.0) * (r * r - s2), 0.0)
    if name == "C2DTLZ2":
        lo = (((f - 1.0) ** 2 + s2 - f ** 2) / (r * r) - 1.0).min(0)
        centre = ((f - 1.0 / math.sqrt(m)) ** 2).sum(0) / (r * r) - 1.0
        return np.maximum(np.minimum(lo, centre), 0.0)
    if name == "C2ConvexDTLZ2":
        mean = f.mean(0)
        return np.maximum(1.0 - (s2 - m * mean ** 2) / (r * r), 0.0)
    if name == "C3DTLZ1":
        return np.maximum(f + 1.0 - 2.0 * f.sum(0), 0.0).sum(0)
    return np.maximum(0.75 * f ** 2 - s2 + 1.0, 0.0).sum(0)


@pytest.mark.parametrize("name", sorted(CONSTRAINED))
@pytest.mark.parametrize("m", [3, 5, 8])
def test_constrained_dtlz_kernels(name, m):
    kind, base = CONSTRAINED[name]
    rng = np.random.default_rng(m * 31 + kind)
    n, d = 700, m + 6
    x = rng.random((n, d), dtype=np.float32)
    # near the Pareto set, so that both feasible and infeasible rows occur
    x[:, m - 1:] = (0.5 + 0.05 * (rng.random((n, d - m + 1)) - 0.5)).astype(np.float32)
    print(f"constrained {name} m={m}", flush=True)
    xt = torch.from_numpy(x).to(DEV)
    f, cv = K.eval_dtlz(kind, xt, m)
    fb, _ = K.eval_dtlz(base, xt, m)
    assert torch.equal(f, fb)                      # the base problem's objectives, bit for bit
    f64 = f.cpu().numpy()
    ref = cv_fp64(name, f64)
    scale = 1.0 + (f64.astype(np.float64) ** 2).sum(0) + np.abs(f64).sum(0)
    got = cv.cpu().numpy()
    assert (got >= 0).all()
    bad = np.abs(got - ref) > 1e-5 * scale
    assert not bad.any(), (name, m, got[bad][:5], ref[bad][:5])


def test_constrained_dtlz_runs_native():
    for name in ("C1DTLZ1", "C2DTLZ2", "C3DTLZ4"):
        prob = getattr(cm, name)(dimension=12, objectives=3)
        assert prob.constraints == 1
        alg = cm.NSGA3(population_size=92, max_generations=10, problem=prob, seed=2, print_progress=False)
        alg.initialize()
        assert alg._args.problem.kind == _lib.DTLZ_KINDS[name]
        res = alg.run()
        f = res.objectives.cpu().numpy().T
        ref = cv_fp64(name, f)
        scale = 1.0 + (f.astype(np.float64) ** 2).sum(0) + np.abs(f).sum(0)
        assert (np.abs(res.constraints.cpu().numpy() - ref) <= 1e-5 * scale).all()


def test_constrained_dtlz_rejects_undefined_objective_counts():
    alg = cm.NSGA3(population_size=92, max_generations=3, problem=cm.C2ConvexDTLZ2(dimension=12, objectives=2),
                   print_progress=False)
    with pytest.raises((ValueError, RuntimeError), match="objective count"):
        alg.initialize()


@pytest.mark.parametrize("kind", ["NSGA3", "RVEA"])
def test_constraint_activation_schedule(kind):
    """constraint_activation_ratio = 0.5 over 10 generations: evaluations of
    generations g with (g + 1) / 10 < 0.5 carry cv = 0; at g = 4 the parents'
    violations are recomputed, and from then on every violation is the
    problem's (dtlz_kernels.cu:1295-1394)."""
    prob = cm.C1DTLZ1(dimension=12, objectives=3, constraint_activation_ratio=0.5)
    alg = cm.Algorithm(kind, population_size=92 if kind == "NSGA3" else 100, max_generations=10, problem=prob,
                       seed=4, print_progress=False)
    alg.initialize()
    print(f"activation {kind}", flush=True)

    def rows():
        # RVEA fills the slots past its active count with placeholders (cv 1e15)
        return int(alg._active.item()) if kind == "RVEA" else 92

    def cv_now():
        return alg.population.constraints.cpu().numpy()[: rows()]

    def true_cv():
        x = alg.population.variables.contiguous()
        return K.eval_dtlz(9, x, 3)[1].cpu().numpy()[: rows()]

    assert not cv_now().any() and true_cv().any()
    for _ in range(4):
        alg.step()
    assert not cv_now().any()                                         # generations 0-3: off
    alg.step()                                                        # generation 4: on
    np.testing.assert_array_equal(cv_now(), true_cv())
    if kind == "RVEA":
        # rows past the live count keep their placeholders through the switch
        pop, live = alg.population, rows()
        assert bool((pop.objectives[live:] == 1e15).all()) and bool((pop.constraints[live:] == 1e15).all())
    alg.run()
    np.testing.assert_array_equal(cv_now(), true_cv())
    # the schedule restarts with reset()
    alg.reset()
    assert not cv_now().any()


def csdp_fp64(x):
    """The car side-impact design problem as dtlz_kernels.cu:719-738 and
    1046-1085 define it: 3 objectives and the sum of 10 violations."""
    x1, x2, x3, x4, x5, x6, x7 = x.astype(np.float64).T
    F = 4.72 - 0.5 * x4 - 0.19 * x2 * x3
    vmbp = 10.58 - 0.674 * x1 * x2 - 0.67275 * x2
    vfd = 16.45 - 0.489 * x3 * x7 - 0.843 * x5 * x6
    f = np.stack([1.98 + 4.9 * x1 + 6.67 * x2 + 6.98 * x3 + 4.01 * x4 + 1.78 * x5 + 0.00001 * x6 + 2.73 * x7,
                  F, 0.5 * (vmbp + vfd)])
    g = [1.16 - 0.3717 * x2 * x4 - 0.0092928 * x3 - 1.0,
         0.261 - 0.0159 * x1 * x2 - 0.06486 * x1 - 0.019 * x2 * x7 + 0.0144 * x3 * x5 + 0.0154464 * x6 - 0.32,
         0.214 + 0.00817 * x5 - 0.045195 * x1 - 0.0135168 * x1 + 0.03099 * x2 * x6 - 0.018 * x2 * x7
         + 0.007176 * x3 + 0.023232 * x3 - 0.00364 * x5 * x6 - 0.018 * x2 * x2 - 0.32,
         0.74 - 0.61 * x2 - 0.031296 * x3 - 0.031872 * x7 + 0.227 * x2 * x2 - 0.32,
         28.98 + 3.818 * x3 - 4.2 * x1 * x2 + 1.27296 * x6 - 2.68065 * x7 - 32.0,
         33.86 + 2.95 * x3 - 5.057 * x1 * x2 - 3.795 * x2 - 3.4431 * x7 + 1.45728 - 32.0,
         46.36 - 9.9 * x2 - 4.4505 * x1 - 32.0,
         F - 4.0, vmbp - 9.9, vfd - 15.7]
    return f, sum(np.maximum(v, 0.0) for v in g)


def test_csdp_kernel_and_run():
    rng = np.random.default_rng(17)
    x = rng.random((3000, 7), dtype=np.float32)
    f, cv = K.eval_dtlz(15, torch.from_numpy(x).to(DEV), 3)
    f64, cv64 = csdp_fp64(x)
    np.testing.assert_allclose(f.cpu().numpy(), f64, rtol=2e-5, atol=2e-4)
    np.testing.assert_allclose(cv.cpu().numpy(), cv64, rtol=2e-5, atol=2e-4)
    assert (cv64 > 0).any()                          # the sample holds infeasible rows
    alg = cm.NSGA3(population_size=92, max_generations=10, problem=cm.CSDP(7, 3), seed=5, print_progress=False)
    alg.initialize()
    assert alg._args.problem.kind == _lib.DTLZ_KINDS["CSDP"]
    res = alg.run()
    f64, cv64 = csdp_fp64(res.variables.cpu().numpy())
    np.testing.assert_allclose(res.objectives.cpu().numpy().T, f64, rtol=2e-5, atol=2e-4)
    np.testing.assert_allclose(res.constraints.cpu().numpy(), cv64, rtol=2e-5, atol=2e-4)
    bad = cm.NSGA3(population_size=92, max_generations=3, problem=cm.CSDP(12, 3), print_progress=False)
    with pytest.raises((ValueError, RuntimeError), match="7 variables"):
        bad.initialize()


def test_activation_switch_prepares_a_python_selector():
    """algorithm.cu:233-241: when the constraints switch on, the parents'
    violations change and a Python selector's prepare() runs again on them."""
    class Keep(cm.PythonEnvironmentSelector):
        def __init__(self):
            self.prepared = []

        def prepare(self, population, context):
            self.prepared.append((context["generation"], float(population["constraints"].abs().sum())))
            return None

        def select(self, parents, offspring, context):
            n = parents["variables"].shape[0]
            return {"indices": torch.arange(n, device=parents["variables"].device)}

    sel = Keep()
    prob = cm.C1DTLZ1(dimension=12, objectives=3, constraint_activation_ratio=0.5)
    alg = cm.NSGA3(population_size=92, max_generations=10, problem=prob, mating=cm.RandomMating(),
                   environment_selector=sel, seed=4, print_progress=False)
    alg.run()
    assert [g for g, _ in sel.prepared] == [0, 4]
    assert sel.prepared[0][1] == 0.0 and sel.prepared[1][1] > 0.0    # violations off, then on
    # ratio 0: the constraints count from the start, and generation 0 is the
    # switch node (is_cv_activation_switch_node(0) = is_cv_active(0)): the
    # selector prepares at initialization and again at generation 0
    sel0 = Keep()
    alg = cm.NSGA3(population_size=92, max_generations=3,
                   problem=cm.C1DTLZ1(dimension=12, objectives=3, constraint_activation_ratio=0.0),
                   mating=cm.RandomMating(), environment_selector=sel0, seed=4, print_progress=False)
    alg.run()
    assert [g for g, _ in sel0.prepared] == [0, 0]
    assert sel0.prepared[0][1] > 0.0 and sel0.prepared[0][1] == sel0.prepared[1][1]


@pytest.mark.parametrize("n,d,m", [(300, 12, 3), (1000, 500, 3), (77, 40, 8), (64, 3, 3)])
def test_dtlz1_kernels_match_oracle_bitexact(n, d, m):
    """DTLZ1 and C1-DTLZ1 evaluated on the GPU equal the oracle's float32
    restatement of the same summation order, bit for bit."""
    import oracle as O
    rng = np.random.default_rng(n + m)
    x = rng.random((n, d), dtype=np.float32)
    x[: n // 2, m - 1:] = (0.5 + 0.05 * (rng.random((n // 2, d - m + 1)) - 0.5)).astype(np.float32)
    print(f"dtlz1 bit-exact n={n} d={d} m={m}", flush=True)
    xt = torch.from_numpy(x).to(DEV)
    for kind, constrained in ((1, False), (9, True)):
        f, cv = K.eval_dtlz(kind, xt, m)
        fo, cvo = O.eval_dtlz1(x, m, constrained=constrained)
        np.testing.assert_array_equal(f.cpu().numpy(), fo)
        np.testing.assert_array_equal(cv.cpu().numpy(), cvo)


@pytest.mark.parametrize("name,ratio,gens,seed", [("DTLZ1", 0.0, 15, 0), ("C1DTLZ1", 0.0, 15, 1),
                                                  ("C1DTLZ1", 0.5, 20, 2), ("C1DTLZ1", 0.9, 10, 3)])
def test_dtlz1_run_matches_oracle(name, ratio, gens, seed):
    """NSGA-III trajectories on DTLZ1 and on C1-DTLZ1 with the constraint
    activation ratio (violations zero until (g + 1) / gens reaches it, then the
    parents are evaluated again and re-sorted) equal the oracle's, bit for bit."""
    import oracle as O
    prob = getattr(cm, name)(dimension=12, objectives=3, constraint_activation_ratio=ratio)
    alg = cm.NSGA3(population_size=92, max_generations=gens, problem=prob, seed=seed, print_progress=False)
    print(f"dtlz1 run {name} ratio={ratio} gens={gens} seed={seed}", flush=True)
    res = alg.run()
    x, f, cv, fr, _ = O.nsga3_run_dtlz1(92, 3, 12, gens, seed, constrained=name != "DTLZ1", ratio=ratio)
    np.testing.assert_array_equal(res.variables.cpu().numpy(), x)
    np.testing.assert_array_equal(res.objectives.cpu().numpy(), f.T)
    np.testing.assert_array_equal(res.constraints.cpu().numpy().reshape(-1), cv)
    np.testing.assert_array_equal(res.auxiliary_indices.cpu().numpy(), fr)


def test_constraint_switch_with_cone_association_matches_oracle():
    """A constraint switch mid-run re-sorts the parents inside the run's
    workspace, and the next selection rebuilds its cone tables (2080
    directions: the cone-pruned association).  C1-DTLZ1, ratio 0.5, ten
    generations, against the oracle, bit for bit."""
    import oracle as O
    prob = cm.C1DTLZ1(dimension=12, objectives=3, constraint_activation_ratio=0.5)
    alg = cm.NSGA3(population_size=2080, max_generations=10, problem=prob, seed=5, print_progress=False)
    print("c1dtlz1 switch with cones n=2080", flush=True)
    res = alg.run()
    assert alg.k >= 2048
    x, f, cv, fr, _ = O.nsga3_run_dtlz1(2080, 3, 12, 10, 5, constrained=True, ratio=0.5)
    np.testing.assert_array_equal(res.variables.cpu().numpy(), x)
    np.testing.assert_array_equal(res.objectives.cpu().numpy(), f.T)
    np.testing.assert_array_equal(res.constraints.cpu().numpy().reshape(-1), cv)
    np.testing.assert_array_equal(res.auxiliary_indices.cpu().numpy(), fr)
