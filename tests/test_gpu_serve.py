"""Served NSGA-III (moea_serve_*, cuda_moea.Server): runs admitted at
different times, sharing steps with runs of other sizes, objectives and
generation counts, each end bit-identical to the same run alone."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402

pytestmark = pytest.mark.gpu

# (objective, n, m, d, partitions, seed, generations)
CASES = [("quad", 92, 3, 12, 12, 0, 6), ("mlp", 64, 5, 30, 6, 1, 9), ("quad", 130, 8, 7, 3, 2, 4),
         ("mlp", 96, 3, 70, 12, 3, 7), ("quad", 256, 2, 100, 20, 4, 11), ("quad", 200, 10, 40, 0, 5, 5),
         ("mlp", 40, 4, 300, 4, 6, 3), ("quad", 2, 3, 5, 12, 7, 8)]


def _run(kind, n, m, d, parts, seed, gens):
    prob = (cm.QuadraticCenters(d, m, seed=seed + 100) if kind == "quad"
            else cm.SeededMLP(d, m, hidden=96 + 32 * seed, seed=seed + 200))
    return cm.NSGA3(population_size=n, max_generations=gens, problem=prob, seed=seed,
                    reference_directions=cm.DasDennisDirections(partitions=parts), print_progress=False)


def _arrays(res):
    return [res.variables.cpu().numpy(), res.objectives.cpu().numpy(), res.constraints.cpu().numpy(),
            res.auxiliary_indices.cpu().numpy()]


def _same(a, b):
    for u, v in zip(_arrays(a), _arrays(b)):
        np.testing.assert_array_equal(u, v)


def _serve(schedule, *, max_instances=8, budget=None):
    """Serve `schedule` [(arrival_step, case)]; runs wait in arrival order when
    no slot or budget is free.  Returns {case index: arrays of its result}."""
    reserved_cfg = dict(max_instances=max_instances, max_rows=2 * 256 * max_instances)
    budget = budget or 64 << 20
    srv = cm.Server(budget, **reserved_cfg)
    pending = sorted(range(len(schedule)), key=lambda i: schedule[i][0])
    owner, out, step = {}, {}, 0
    while len(out) < len(schedule):
        while pending and schedule[pending[0]][0] <= step:
            i = pending[0]
            slot = srv.admit(_run(*schedule[i][1]))
            if slot is None:
                break
            owner[slot] = i
            pending.pop(0)
        for slot in srv.step():
            i = owner.pop(slot)
            assert srv.generation(slot) == schedule[i][1][-1]
            out[i] = _arrays(srv.result(slot))
            srv.release(slot)
        step += 1
        assert step < 1000
    return out


def test_serve_all_at_once_equals_single_runs():
    print("serve all at once", flush=True)
    out = _serve([(0, c) for c in CASES])
    for i, c in enumerate(CASES):
        print(f"  case {c}", flush=True)
        for u, v in zip(out[i], _arrays(_run(*c).run())):
            np.testing.assert_array_equal(u, v)


def test_serve_staggered_with_slot_reuse_equals_single_runs():
    print("serve staggered, 3 slots", flush=True)
    arrivals = [0, 0, 1, 2, 2, 5, 6, 9]
    out = _serve(list(zip(arrivals, CASES)), max_instances=3)
    for i, c in enumerate(CASES):
        print(f"  case {c}", flush=True)
        for u, v in zip(out[i], _arrays(_run(*c).run())):
            np.testing.assert_array_equal(u, v)


def test_serve_alone_equals_mixed():
    print("serve alone vs mixed", flush=True)
    target = 3
    alone = _serve([(0, CASES[target])])[0]
    mixed = _serve([(3, CASES[target])] + [(a, c) for a, c in zip([0, 1, 1, 4], CASES[:3] + CASES[4:5])])[0]
    for u, v in zip(alone, mixed):
        np.testing.assert_array_equal(u, v)


def test_serve_budget_admission_and_release():
    print("serve budget", flush=True)
    a, b = _run(*CASES[4]), _run(*CASES[1])
    cfg = dict(max_instances=4, max_rows=4096)
    probe = cm.Server(64 << 20, **cfg)
    import ctypes
    from cuda_moea import _lib
    reserved = _lib.lib.moea_serve_reserved_bytes(ctypes.byref(probe._cfg))
    del probe
    need_a, need_b = cm.Server.instance_bytes(a), cm.Server.instance_bytes(b)
    assert need_a > 0 and need_b > 0
    srv = cm.Server(reserved + max(need_a, need_b) + 256, **cfg)
    sa = srv.admit(a)
    assert sa is not None
    assert srv.admit(b) is None                 # budget full
    done = []
    while not done:
        done = srv.step()
    assert done == [sa]
    res_a = srv.result(sa)
    srv.release(sa)
    sb = srv.admit(b)                           # the freed space is reused
    assert sb is not None
    while not srv.step():
        pass
    _same(res_a, _run(*CASES[4]).run())
    _same(srv.result(sb), _run(*CASES[1]).run())
    with pytest.raises(ValueError):
        srv.admit(_run("quad", 91, 3, 12, 12, 0, 5))
    with pytest.raises(Exception):
        srv.release(sa + 1 if sa + 1 != sb else sa + 2)


def test_served_quadratic_runs_match_oracle():
    """Every quadratic-center run served in a staggered mix (slots reused,
    MLP runs beside them) ends exactly where the oracle's restatement of the
    reference NSGA-III loop ends: variables, objectives and fronts."""
    import oracle as O
    print("serve staggered vs oracle", flush=True)
    arrivals = [0, 0, 1, 2, 2, 5, 6, 9]
    out = _serve(list(zip(arrivals, CASES)), max_instances=3)
    for i, (kind, n, m, d, parts, seed, gens) in enumerate(CASES):
        if kind != "quad":
            continue
        print(f"  oracle case n={n} m={m} d={d} parts={parts} seed={seed} gens={gens}", flush=True)
        prob = cm.QuadraticCenters(d, m, seed=seed + 100)
        bounds = np.tile(np.array([0.0, 1.0], np.float32), d)
        x, f, cv, fr, _ = O.nsga3_run(n, m, d, gens, seed, bounds, prob.centers.numpy(), partitions=parts)
        np.testing.assert_array_equal(out[i][0], x)
        np.testing.assert_array_equal(out[i][1], f.T)
        np.testing.assert_array_equal(out[i][2], cv)
        np.testing.assert_array_equal(out[i][3], fr)


def test_serve_full_size_instance_alone_equals_mixed():
    """The largest served shape (N = 8192, M = 10, D = 4096, seeded MLP
    D -> 256 -> 10, directions sized from N) served beside small runs that
    arrive and leave around it equals the same run served alone."""
    big = ("mlp_big", 8192, 10, 4096, 0, 9, 3)

    def make(case):
        if case[0] == "mlp_big":
            _, n, m, d, parts, seed, gens = case
            prob = cm.SeededMLP(d, m, hidden=256, seed=seed + 200, device=torch.device("cuda", 0))
            return cm.NSGA3(population_size=n, max_generations=gens, problem=prob, seed=seed,
                            print_progress=False)
        return _run(*case)

    def serve(schedule):
        srv = cm.Server(8 << 30, max_instances=6, max_rows=2 * 8192 + 2 * 256 * 5)
        pending = sorted(range(len(schedule)), key=lambda i: schedule[i][0])
        owner, out, step = {}, {}, 0
        while len(out) < len(schedule):
            while pending and schedule[pending[0]][0] <= step:
                slot = srv.admit(make(schedule[pending[0]][1]))
                assert slot is not None
                owner[slot] = pending.pop(0)
            for slot in srv.step():
                i = owner.pop(slot)
                out[i] = _arrays(srv.result(slot))
                srv.release(slot)
            step += 1
            assert step < 100
        return out

    print("serve full-size instance alone", flush=True)
    alone = serve([(0, big)])[0]
    print("serve full-size instance mixed", flush=True)
    mixed = serve([(1, big), (0, CASES[0]), (0, CASES[1]), (2, CASES[2]), (3, CASES[3])])[0]
    for u, v in zip(alone, mixed):
        np.testing.assert_array_equal(u, v)
    f = alone[1]
    assert f.shape == (8192, 10) and np.all(np.isfinite(f))
