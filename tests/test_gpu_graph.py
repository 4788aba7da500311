"""Generation graphs and the peel-plan autotuner of the batched engine.

A batch with graph=True replays one captured HIP graph per generation
(moea_batch_set_graph).  Every replay must equal the direct launches bit for
bit: the populations are compared after every generation, with eager torch
work and synchronizations between the replays, and the captured graph must
hold kernel nodes only (no memset nodes).  The peel plans the autotuner picks
from must not change any result, and a tuned run must equal an untuned one."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402
from cuda_moea import _lib  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)

# (n, m, d, partitions, seed) quadratic-center and (n, m, d, partitions, seed, hidden) MLP runs
QUAD = [(92, 3, 12, 12, 0), (64, 5, 30, 6, 1), (130, 8, 7, 3, 2), (256, 2, 100, 20, 3),
        (1000, 4, 50, 0, 5), (2, 3, 5, 12, 6), (4096, 3, 20, 12, 7)]
MLP = [(96, 3, 40, 12, 8, 256), (64, 5, 130, 6, 9, 100)]


def _runs(gens):
    runs = [cm.NSGA3(population_size=n, max_generations=gens, problem=cm.QuadraticCenters(d, m, seed=s + 100),
                     seed=s, reference_directions=cm.DasDennisDirections(partitions=p), print_progress=False)
            for n, m, d, p, s in QUAD]
    runs += [cm.NSGA3(population_size=n, max_generations=gens, problem=cm.SeededMLP(d, m, hidden=h, seed=s + 200),
                      seed=s, reference_directions=cm.DasDennisDirections(partitions=p), print_progress=False)
             for n, m, d, p, s, h in MLP]
    return runs


def _arrays(res):
    return [res.variables.cpu().numpy(), res.objectives.cpu().numpy(), res.constraints.cpu().numpy(),
            res.auxiliary_indices.cpu().numpy()]


def _same(a, b):
    for u, v in zip(_arrays(a), _arrays(b)):
        np.testing.assert_array_equal(u, v)


def test_graph_replays_equal_direct_launches_every_generation():
    gens = 9
    direct = cm.BatchNSGA3(_runs(gens))
    graph = cm.BatchNSGA3(_runs(gens), graph=True)
    direct.initialize()
    graph.initialize()
    scratch = torch.zeros(1 << 20, device=DEV)
    for g in range(gens):
        print(f"graph vs direct, generation {g + 1}", flush=True)
        direct.step()
        graph.step()
        # eager work of the program's own between the replays
        scratch.add_(1.0).mul_(0.5)
        torch.cuda.synchronize()
        for i in range(len(direct.runs)):
            _same(graph.result(i), direct.result(i))
    info = graph.graph_info()
    print(f"graph info {info}", flush=True)
    assert info["memset_nodes"] == 0
    assert info["kernel_nodes"] == info["nodes"] > 0
    assert info["captures"] == 1                       # the set of runs never changed
    assert direct.graph_info()["captures"] == 0


def test_every_peel_plan_gives_the_same_runs():
    gens = 6
    ref = cm.BatchNSGA3(_runs(gens)).run()
    for div, rows in [(1, 32), (2, 16), (2, 32), (4, 8)]:
        print(f"peel plan ({div}, {rows})", flush=True)
        b = cm.BatchNSGA3(_runs(gens), graph=True)
        b.initialize()
        _lib.check(_lib.lib.moea_batch_set_peel_plan(b._handle, div, rows))
        for res, want in zip(b.run(), ref):
            _same(res, want)


def test_autotuned_run_equals_untuned_run():
    gens = 24
    untuned = cm.BatchNSGA3(_runs(gens)).run()
    tuned = cm.BatchNSGA3(_runs(gens), graph=True)
    tuned.initialize()
    print("autotune, 2 generations per visit", flush=True)
    report = tuned.autotune(trials=2)
    print(f"autotune report {report}", flush=True)
    assert report["generations"] == 16 and tuned.generation == 16
    assert 0 <= report["chosen"] < 4
    timed = [ms for ms in report["ms_per_generation"] if ms >= 0]
    assert timed and min(timed) == report["ms_per_generation"][report["chosen"]]
    for res, want in zip(tuned.run(), untuned):
        _same(res, want)
    with pytest.raises(ValueError):
        tuned.autotune(trials=1)


def test_served_graph_equals_served_direct():
    """Runs arriving and leaving (so the graph is recaptured on each change)
    give the same results with and without generation graphs."""
    cases = [((n, m, d, p, s), g) for (n, m, d, p, s), g in zip(QUAD, [4, 7, 3, 6, 5, 8, 4])]

    def serve(graph):
        srv = cm.Server(256 << 20, max_instances=4, max_rows=2 * 4096 + 2 * 1000 * 3, graph=graph)
        arrivals = [0, 0, 1, 2, 3, 3, 6]
        pending = list(range(len(cases)))
        owner, out, step = {}, {}, 0
        while len(out) < len(cases):
            while pending and arrivals[pending[0]] <= step:
                (n, m, d, p, s), g = cases[pending[0]]
                run = cm.NSGA3(population_size=n, max_generations=g, problem=cm.QuadraticCenters(d, m, seed=s + 100),
                               seed=s, reference_directions=cm.DasDennisDirections(partitions=p), print_progress=False)
                slot = srv.admit(run)
                if slot is None:
                    break
                owner[slot] = pending.pop(0)
            for slot in srv.step():
                out[owner.pop(slot)] = _arrays(srv.result(slot))
                srv.release(slot)
            step += 1
            assert step < 200
        return out

    print("serve with graphs", flush=True)
    a = serve(True)
    print("serve direct", flush=True)
    b = serve(False)
    for i in range(len(cases)):
        for u, v in zip(a[i], b[i]):
            np.testing.assert_array_equal(u, v)


def _hip():
    """The HIP runtime this process already loaded (torch's and the
    library's), for the graph API calls of the test."""
    import ctypes
    with open("/proc/self/maps") as f:
        paths = sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln and "/" in ln})
    assert paths, "HIP runtime not loaded"
    hip = ctypes.CDLL(paths[0])
    vp = ctypes.c_void_p
    for name, args in (("hipStreamCreateWithFlags", [ctypes.POINTER(vp), ctypes.c_uint]),
                       ("hipStreamBeginCapture", [vp, ctypes.c_int]),
                       ("hipStreamEndCapture", [vp, ctypes.POINTER(vp)]),
                       ("hipGraphGetNodes", [vp, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_size_t)]),
                       ("hipGraphNodeGetType", [vp, ctypes.POINTER(ctypes.c_int)]),
                       ("hipGraphInstantiate", [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_size_t]),
                       ("hipGraphLaunch", [vp, vp]), ("hipStreamSynchronize", [vp]),
                       ("hipGraphExecDestroy", [vp]), ("hipGraphDestroy", [vp]), ("hipStreamDestroy", [vp])):
        fn = getattr(hip, name)
        fn.restype, fn.argtypes = ctypes.c_int, args
    return hip


def test_single_population_generation_replays_from_a_graph():
    """One NSGA-III generation through the C ABI (moea_reproduce,
    moea_eval_quadratic, moea_nsga3_select, whose sort includes the
    cooperative front-peeling launch), captured from a stream into a graph:
    only kernel nodes, and every replay equals the direct calls."""
    import ctypes
    from cuda_moea import kernels as K
    hip = _hip()
    n, m, d = 2048, 3, 30
    v = torch.from_numpy(K.das_dennis(n, m)).to(DEV)
    k = v.shape[1]
    bounds = torch.tensor([0.0, 1.0] * d, device=DEV)
    centers = torch.rand(m, d, generator=torch.Generator().manual_seed(3)).to(DEV)
    par = K.Population(K.init_population(n, d, bounds, 11), None, None)
    par.f, par.cv = K.eval_quadratic(par.x, centers)
    ideal0 = K.ideal_point(par.f)
    off = K.Population.empty(n, d, m, DEV)
    nxt = K.Population.empty(n, d, m, DEV)
    idx = torch.empty(n, dtype=torch.int32, device=DEV)
    nfr = torch.empty(n, dtype=torch.int32, device=DEV)
    stats = torch.zeros(8, dtype=torch.int32, device=DEV)
    ideal = ideal0.clone()
    ws = K.workspace(_lib.lib.moea_nsga3_workspace_bytes(n, m, k), DEV)
    torch.cuda.synchronize()

    def generation(stream):
        r = _lib.ReproduceArgs()
        r.n, r.d, r.active, r.mating = n, d, n, _lib.MATING_RANDOM
        r.parents_x, r.bounds = par.x.data_ptr(), bounds.data_ptr()
        r.sbx_eta, r.sbx_prob, r.sbx_copy_prob = 30.0, 1.0, 0.5
        r.mutation, r.pm_eta, r.pm_prob = _lib.MUTATION_POLYNOMIAL, 20.0, 1.0
        r.seed, r.generation = 5, 7
        r.offspring_x, r.parent_indices, r.stream = off.x.data_ptr(), idx.data_ptr(), stream
        _lib.check(_lib.lib.moea_reproduce(ctypes.byref(r)))
        _lib.check(_lib.lib.moea_eval_quadratic(off.x.data_ptr(), n, d, centers.data_ptr(), m, off.f.data_ptr(), n,
                                                off.cv.data_ptr(), stream))
        a = _lib.NSGA3SelectArgs()
        a.n, a.m, a.d, a.k = n, m, d, k
        a.parents, a.offspring, a.next = par.c(), off.c(), nxt.c()
        a.next_fronts, a.directions, a.ideal = nfr.data_ptr(), v.data_ptr(), ideal.data_ptr()
        a.seed, a.generation = 5, 7
        a.stats, a.workspace, a.workspace_bytes, a.stream = stats.data_ptr(), ws.data_ptr(), ws.numel(), stream
        _lib.check(_lib.lib.moea_nsga3_select(ctypes.byref(a)))

    s = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(s), 1) == 0
    print("single generation: direct", flush=True)
    generation(s.value)
    assert hip.hipStreamSynchronize(s) == 0
    want = [t.clone() for t in (off.x, off.f, nxt.x, nxt.f, nxt.cv, nfr, stats[:3])]
    ideal.copy_(ideal0)
    torch.cuda.synchronize()
    print("single generation: capture", flush=True)
    assert hip.hipStreamBeginCapture(s, 1) == 0
    generation(s.value)
    g = ctypes.c_void_p()
    assert hip.hipStreamEndCapture(s, ctypes.byref(g)) == 0 and g.value
    count = ctypes.c_size_t(0)
    assert hip.hipGraphGetNodes(g, None, ctypes.byref(count)) == 0 and count.value > 0
    nodes = (ctypes.c_void_p * count.value)()
    assert hip.hipGraphGetNodes(g, nodes, ctypes.byref(count)) == 0
    types = []
    for nd in nodes:
        t = ctypes.c_int(-1)
        assert hip.hipGraphNodeGetType(ctypes.c_void_p(nd), ctypes.byref(t)) == 0
        types.append(t.value)
    print(f"  {len(types)} nodes, types {sorted(set(types))}", flush=True)
    assert set(types) == {0}                       # hipGraphNodeTypeKernel only: no memset nodes
    x = ctypes.c_void_p()
    assert hip.hipGraphInstantiate(ctypes.byref(x), g, None, None, 0) == 0
    scratch = torch.zeros(1 << 18, device=DEV)
    try:
        for rep in range(3):
            print(f"single generation: replay {rep}", flush=True)
            for t in (off.x, off.f, nxt.x, nxt.f, nxt.cv):
                t.fill_(-1.0)
            nfr.fill_(-9)
            scratch.add_(1.0)
            torch.cuda.synchronize()
            assert hip.hipGraphLaunch(x, s) == 0
            assert hip.hipStreamSynchronize(s) == 0
            for a_, b_ in zip([off.x, off.f, nxt.x, nxt.f, nxt.cv, nfr, stats[:3]], want):
                assert torch.equal(a_, b_)
    finally:
        hip.hipGraphExecDestroy(x)
        hip.hipGraphDestroy(g)
        hip.hipStreamDestroy(s)
