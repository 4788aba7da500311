"""Headline benchmark: NSGA-III generations per second at N=32768, M=3, D=1024
(quadratic-center objectives), one MI355X.

A step is one generation (mating -> SBX -> PM -> objective evaluation ->
NSGA-III selection) of the native run loop, inputs resident in HBM.  W warmup
generations, then K timed generations between two synchronized HIP events on
the run stream.  An identical second run (same seeds, so the same data) times
the same generations with HIP events around every stage on the run stream;
`roofline` covers the kernel with the largest average launch time there
(front peeling: algorithmic bytes against 8 TB/s; dominance: unordered pairs x
M VALU lane-ops against 256 CUs x 64 lanes x 2.4 GHz), the other beside it in
`roofline_other`.  The CPU baseline times the first generations of the oracle
(C restatement of the reference algorithm) on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cuda-moea_amd"))

N, M, D = 32768, 3, 1024
VALU_PEAK_CMP = 256 * 64 * 2.4e9          # VALU lane-ops / s (256 CUs x 64 lanes x 2.4 GHz)
HBM_PEAK = 8.0e12


def _traffic_from_profiles(kernel: str) -> float | None:
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, each scaled by
    what the counter reads for a copy kernel of known size) from the committed
    rocprofv3 PMC summary of the generations 0-2 launches."""
    p = ROOT / "profiles" / "nds_n65536_m3_pmc_traffic.md"
    if not p.exists():
        return None
    key = f"traffic_bytes_per_launch[{kernel}]:"
    for line in p.read_text().splitlines():
        if line.startswith(key):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def cpu_baseline(threads: int) -> dict:
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import torch
    import oracle as O
    g = torch.Generator().manual_seed(0)
    centers = torch.rand(M, D, generator=g).numpy()
    bounds = np.tile(np.array([0.0, 1.0], np.float32), D)
    # the first `gens` generations of the full NSGA-III loop at the benchmark shape
    gens = 5
    _, _, _, _, ms = O.nsga3_run(N, M, D, 100, 0, bounds, centers, limit=gens)
    return {"value": 1000.0 * gens / ms, "unit": "gens/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"generations 0-{gens - 1} of the oracle NSGA-III loop at N={N} M={M} D={D} "
                      f"(quadratic-center, K=32640), {O.num_threads()} OpenMP threads, "
                      f"{ms / 1000:.1f} s"}


def config0_cpu(threads: int) -> dict:
    """Config 0 on the oracle (the C restatement; the reference has no CPU
    path): NSGA-III N = 92, M = 3, H = 12 (K = 91), D = 12, quadratic-center
    objectives with centers from torch.rand after manual_seed(0), SBX eta 30
    p 1, PM eta 20 rate 1/D, tournament mating, 200 generations, seed 0;
    timed single-threaded and on `threads` OpenMP threads."""
    import platform
    import numpy as np
    import torch
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    g = torch.Generator().manual_seed(0)
    centers = torch.rand(3, 12, generator=g).numpy()
    bounds = np.tile(np.array([0.0, 1.0], np.float32), 12)
    out = {}
    for t in (1, threads):
        O.set_threads(t)
        _, _, _, _, ms = O.nsga3_run(92, 3, 12, 200, 0, bounds, centers, partitions=12)
        out[f"ms_{t}_threads"] = round(ms, 2)
    O.set_threads(threads)
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    out.update({"workload": "config 0: NSGA-III N=92 M=3 K=91 D=12, 200 generations, oracle on the host",
                "nproc": os.cpu_count(), "cpu_model": cpu, "machine": platform.machine()})
    return out


def batched_config(generations: int, count: int, plain_sample: int) -> dict:
    """Config 3: `count` independent NSGA-III runs (N log-uniform in [64, 4096],
    even; M in {3, 5, 8} with Das-Dennis H = 12 / 6 / 3; D log-uniform in
    [10, 1000]; per-run seeded quadratic centers) advanced together by one
    BatchNSGA3, against the plain one-run-at-a-time GPU path (cuda_moea.NSGA3)
    timed on the first `plain_sample` runs.  The batch is timed three ways on
    the same runs: direct launches, one replayed hipGraph per generation, and
    graphs with the peel plan autotuned on the runs' own first generations (the
    configured path, which `batched_runs_per_s` reports; the tuning
    generations are inside the timing).  Runs/s count whole runs of
    `generations` generations; the initial evaluation is outside the timing,
    as in Result.total_ms."""
    import math
    import numpy as np
    import torch
    import cuda_moea as cm
    rng = np.random.default_rng(2024)
    parts = {3: 12, 5: 6, 8: 3}
    specs = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(64), math.log(4096)))))
        n += n & 1
        m = int(rng.choice([3, 5, 8]))
        d = int(round(math.exp(rng.uniform(math.log(10), math.log(1000)))))
        specs.append((n, m, d))

    def make(i):
        n, m, d = specs[i]
        return cm.NSGA3(population_size=n, max_generations=generations,
                        problem=cm.QuadraticCenters(d, m, seed=10_000 + i), seed=i,
                        reference_directions=cm.DasDennisDirections(partitions=parts[m]),
                        print_progress=False)

    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)

    def timed_batch(graph: bool, tune: bool):
        batch = cm.BatchNSGA3([make(i) for i in range(count)], graph=graph)
        batch.initialize()
        stream.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(stream)
        report = batch.autotune(trials=3) if tune else None      # the tuning generations are the runs' own
        batch.step(generations - batch.generation)
        t1.record(stream)
        t1.synchronize()
        first = batch.result(0).objectives.cpu()
        info = batch.graph_info()
        del batch
        torch.cuda.empty_cache()
        return t0.elapsed_time(t1) / 1e3, report, info, first

    direct_s, _, _, f_direct = timed_batch(False, False)
    graph_s, _, graph_info, f_graph = timed_batch(True, False)
    tuned_s, tune, _, f_tuned = timed_batch(True, True)
    if not (torch.equal(f_direct, f_graph) and torch.equal(f_direct, f_tuned)):
        raise RuntimeError("batched paths disagree: direct, graph and tuned results differ")
    # the configured path: one captured graph per generation, peel plan autotuned
    batched_s = tuned_s
    plain_s = 0.0
    for i in range(plain_sample):
        alg = make(i)
        alg.initialize()
        stream.synchronize()
        a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a0.record(stream)
        alg._call(cm.algorithm.lib.moea_run_step, generations)
        a1.record(stream)
        a1.synchronize()
        plain_s += a0.elapsed_time(a1) / 1e3
    batched_rps = count / batched_s
    plain_rps = plain_sample / plain_s
    return {"workload": "config 3: NSGA-III, N log-uniform [64,4096], M in {3,5,8}, D log-uniform [10,1000], "
                        "quadratic-center",
            "instances": count, "generations": generations,
            "path": "hipGraph replay per generation, peel plan autotuned on the runs' own first 24 generations",
            "batched_runs_per_s": round(batched_rps, 2), "batched_s": round(batched_s, 3),
            "ms_per_batched_generation": round(batched_s / generations * 1e3, 4),
            "ms_per_batched_generation_direct": round(direct_s / generations * 1e3, 4),
            "ms_per_batched_generation_graph": round(graph_s / generations * 1e3, 4),
            "graph_kernel_nodes": graph_info["kernel_nodes"], "graph_memset_nodes": graph_info["memset_nodes"],
            "autotune": {"chosen_plan": list(tune["plan"]),
                         "ms_per_generation": [round(x, 4) for x in tune["ms_per_generation"]]},
            "plain_runs_per_s": round(plain_rps, 3), "plain_sample_runs": plain_sample,
            "speedup_vs_plain": round(batched_rps / plain_rps, 2),
            "total_individuals": sum(s[0] for s in specs)}


def serving_config(seconds: float, rate: float, plain_sample: int, budget_bytes: int) -> dict:
    """Config 4: a Poisson stream of NSGA-III run requests at `rate`/s for
    `seconds` s (N log-uniform in [64, 8192], even; M in {3, 5, 8, 10}; D
    log-uniform in [10, 4096]; G uniform in [50, 500]; quadratic-center or
    seeded-MLP D->256->M objectives 50/50; directions sized from N) served by
    one cuda_moea.Server whose workspace is the HBM budget.  The loop admits
    arrived requests, runs one batched generation of every live run, waits for
    it, and retires the runs that finished; a request's latency runs from its
    arrival time to the end of the step that finished it.  One-at-a-time
    serving is the same arrivals through a FIFO queue with one plain run at a
    time: each request's service time is its G times a per-generation time
    drawn from `plain_sample` requests of the stream run alone end to end
    (cuda_moea.NSGA3.run, allocation and initialization included)."""
    import collections
    import math
    import numpy as np
    import torch
    import cuda_moea as cm
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    rng = np.random.default_rng(777)
    arrivals = []
    t = rng.exponential(1.0 / rate)
    while t < seconds:
        arrivals.append(t)
        t += rng.exponential(1.0 / rate)
    total = len(arrivals)
    specs = []
    for i in range(total):
        n = int(round(math.exp(rng.uniform(math.log(64), math.log(8192)))))
        n += n & 1
        m = int(rng.choice([3, 5, 8, 10]))
        d = int(round(math.exp(rng.uniform(math.log(10), math.log(4096)))))
        g = int(rng.integers(50, 501))
        specs.append((n, m, d, g, bool(rng.integers(0, 2))))
    # request payloads are resident in HBM before the stream starts
    problems = [cm.SeededMLP(d, m, hidden=256, seed=50_000 + i, device=dev) if mlp
                else cm.QuadraticCenters(d, m, seed=50_000 + i)
                for i, (n, m, d, g, mlp) in enumerate(specs)]

    def make(i, problem=None):
        n, m, d, g, _ = specs[i]
        return cm.NSGA3(population_size=n, max_generations=g, problem=problem or problems[i], seed=i,
                        print_progress=False)

    # warm-up: every kernel variant once, outside the measurement
    warm = cm.Server(1 << 30, max_instances=8, max_rows=2 * 8192 * 8)
    for m in (3, 5, 8, 10):
        for mlp in (False, True):
            prob = cm.SeededMLP(64, m, hidden=256, seed=1, device=dev) if mlp else cm.QuadraticCenters(64, m)
            warm.admit(cm.NSGA3(population_size=128, max_generations=2, problem=prob, seed=0,
                                print_progress=False))
    while warm.live:
        for slot in warm.step():
            warm.release(slot)
    stream.synchronize()
    del warm

    from cuda_moea import _lib
    srv = cm.Server(budget_bytes, max_instances=512, max_rows=2 * 8192 * 512)
    _lib.profile_enable(True)
    owner, queue, latency = {}, collections.deque(), []
    in_flight, step_ms = [], []
    nxt, steps = 0, 0
    t0 = time.perf_counter()
    report = 10.0
    while len(latency) < total:
        now = time.perf_counter() - t0
        while nxt < total and arrivals[nxt] <= now:
            queue.append(nxt)
            nxt += 1
        while queue:
            slot = srv.admit(make(queue[0]))
            if slot is None:
                break
            owner[slot] = queue.popleft()
        if not owner:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
            continue
        a = time.perf_counter()
        done = srv.step()
        stream.synchronize()
        b = time.perf_counter()
        step_ms.append((b - a) * 1e3)
        in_flight.append(len(owner))
        steps += 1
        for slot in done:
            latency.append(b - t0 - arrivals[owner.pop(slot)])
            srv.release(slot)
        if b - t0 > report:
            print(f"serving: t={b - t0:.0f}s arrived={nxt}/{total} done={len(latency)} "
                  f"in_flight={len(owner)} queued={len(queue)}", file=sys.stderr, flush=True)
            report += 10.0
    makespan = time.perf_counter() - t0
    stages = _lib.profile_read()
    _lib.profile_enable(False)
    del srv
    torch.cuda.empty_cache()

    # one-at-a-time: per-generation service times of sampled requests run alone
    pick = rng.choice(total, size=min(plain_sample, total), replace=False)
    per_gen = []
    for i in pick:
        stream.synchronize()
        a = time.perf_counter()
        make(int(i)).run()
        stream.synchronize()
        per_gen.append((time.perf_counter() - a) / specs[int(i)][3])
    draw = rng.choice(np.asarray(per_gen), size=total)
    finish, plain_lat = 0.0, []
    for i in range(total):
        start = max(arrivals[i], finish)
        finish = start + draw[i] * specs[i][3]
        plain_lat.append(finish - arrivals[i])
    lat = np.asarray(latency)
    plat = np.asarray(plain_lat)
    return {"workload": f"config 4: Poisson arrivals {rate:g}/s for {seconds:g} s, N log-uniform [64,8192], "
                        "M in {3,5,8,10}, D log-uniform [10,4096], G uniform [50,500], quadratic-center / "
                        "seeded MLP D->256->M 50/50, directions sized from N",
            "requests": total, "hbm_budget_bytes": budget_bytes,
            "runs_per_s": round(total / makespan, 2), "makespan_s": round(makespan, 2),
            "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
            "steps": steps, "mean_step_ms": round(float(np.mean(step_ms)), 3),
            "mean_runs_in_flight": round(float(np.mean(in_flight)), 1),
            "stage_ms_per_step": {k: round(v[0] / steps, 3) for k, v in stages.items() if v[1]},
            "one_at_a_time": {
                "runs_per_s": round(total / finish, 3),
                "p50_latency_s": round(float(np.percentile(plat, 50)), 2),
                "p99_latency_s": round(float(np.percentile(plat, 99)), 2),
                "mean_service_s": round(float(np.mean(draw * np.asarray([s[3] for s in specs]))), 4),
                "sampled_runs": len(per_gen),
                "method": "FIFO queue over the same arrivals; service = G x per-generation time of a sampled "
                          "request run alone (NSGA3.run, wall clock)"},
            "p50_speedup": round(float(np.percentile(plat, 50) / np.percentile(lat, 50)), 1),
            "p99_speedup": round(float(np.percentile(plat, 99) / np.percentile(lat, 99)), 1)}


def selection_bench(calls: int) -> dict:
    """The standalone selection bench of config 1: one NSGA-III selection
    (moea_nsga3_select: constrained NDS, split, normalization, association,
    niching, gather) over 2N = 65536 merged rows s*u (u uniform on the
    positive unit-sphere orthant, s in [1, 2]), all feasible, K = 32640,
    in-kernel draws; `calls` selections between HIP events."""
    import numpy as np
    import torch
    from cuda_moea import kernels as K
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    n, m = N, M
    rng = np.random.default_rng(2024)
    g = np.abs(rng.standard_normal((2 * n, m)))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    fm = np.ascontiguousarray((rng.uniform(1.0, 2.0, (2 * n, 1)) * u).T.astype(np.float32))
    f = torch.from_numpy(fm).to(dev)
    cv = torch.zeros(2 * n, dtype=torch.float32, device=dev)
    x = torch.zeros(2 * n, 2, dtype=torch.float32, device=dev)
    w = torch.from_numpy(K.das_dennis(n, m)).to(dev)
    ideal = f[:, :n].min(dim=1).values.contiguous()
    par = K.Population(x[:n].contiguous(), f[:, :n].contiguous(), cv[:n].contiguous())
    ofs = K.Population(x[n:].contiguous(), f[:, n:].contiguous(), cv[n:].contiguous())
    K.nsga3_select(par, ofs, w, ideal.clone(), seed=1, generation=0)
    stream.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for c in range(calls):
        _, _, st = K.nsga3_select(par, ofs, w, ideal.clone(), seed=1, generation=c)
    t1.record(stream)
    t1.synchronize()
    ms = t0.elapsed_time(t1) / calls
    return {"workload": "NSGA-III selection alone, 2N = 65536 rows s*u on the positive orthant, K = 32640",
            "ms_per_selection": round(ms, 4), "calls": calls, "max_front": int(st[0].item()),
            "n_prior": int(st[1].item()), "n_last": int(st[2].item())}


def rvea_config(generations: int, warmup: int) -> dict:
    """Config 2: RVEA, N = 16384, M = 5, D = 131072, adaptive reference
    directions (frequency 0.1), alpha = 2, seeded-MLP objective D->256->5
    (weights drawn on the GPU), K timed generations after W warmup
    generations of the 200-generation schedule (the APD penalty and the
    annealed operator parameters follow generation / 200)."""
    import torch
    import cuda_moea as cm
    from cuda_moea import _lib
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    alg = cm.RVEA(population_size=16384, max_generations=200,
                  problem=cm.SeededMLP(131072, 5, hidden=256, seed=0, device=dev),
                  environment_selector=cm.RVEAEnvironmentSelector(alpha=2.0),
                  reference_directions=cm.AdaptiveRVEADirections(frequency=0.1),
                  seed=0, print_progress=False)
    alg.initialize()
    if warmup:
        alg._call(_lib.lib.moea_run_step, warmup)
    stream.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    alg._call(_lib.lib.moea_run_step, generations)
    t1.record(stream)
    t1.synchronize()
    ms = t0.elapsed_time(t1) / generations
    _lib.profile_enable(True)
    alg._call(_lib.lib.moea_run_step, 2)
    stream.synchronize()
    stages = _lib.profile_read()
    _lib.profile_enable(False)
    flops = 2.0 * 16384 * 131072 * 256
    ev = stages.get("evaluate", (0.0, 0))
    ev_ms = ev[0] / max(ev[1], 1)
    out = {"workload": "config 2: RVEA N=16384 M=5 D=131072, seeded MLP D->256->5, adaptive directions "
                       "(fr 0.1), alpha 2", "K_directions": alg.k,
           "ms_per_gen": round(ms, 3), "gens_per_s": round(1000.0 / ms, 3),
           "generations_timed": generations, "warmup": warmup,
           "projected_200_generations_s": round(200 * ms / 1000.0, 2),
           "stage_ms_per_generation": {k: round(v[0] / max(v[1], 1), 3) for k, v in stages.items() if v[1]},
           "mlp_flop_per_gen": flops,
           "mlp_tflops": round(flops / (ev_ms / 1e3) / 1e12, 1) if ev_ms > 0 else None}
    del alg
    torch.cuda.empty_cache()
    return out


def dtlz1_comparison(generations: int, warmup: int) -> dict:
    """NSGA-III on DTLZ1, N = 32768, M = 3, D = 500, in the reference's own
    benchmark setting (SBX eta 30, p 1, per-variable copy probability 0.5; PM
    eta 20, rate 1/D; tournament mating; directions sized from N), timed like
    Result.total_ms.  The reference publishes 5.1406 ms/gen for this shape on
    an RTX PRO 6000 Blackwell (population_timing_summary.csv:9)."""
    import torch
    import cuda_moea as cm
    from cuda_moea import _lib
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    alg = cm.NSGA3(population_size=32768, max_generations=warmup + generations,
                   problem=cm.DTLZ1(dimension=500, objectives=3),
                   crossover=cm.SBX(eta_initial=30.0, eta_final=30.0, probability=1.0,
                                    variable_copy_probability=0.5),
                   seed=0, print_progress=False)
    alg.initialize()
    if warmup:
        alg._call(_lib.lib.moea_run_step, warmup)
    stream.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    alg._call(_lib.lib.moea_run_step, generations)
    t1.record(stream)
    t1.synchronize()
    ms = t0.elapsed_time(t1) / generations
    return {"workload": "NSGA-III DTLZ1 N=32768 M=3 D=500 (reference benchmark setting)",
            "ms_per_gen": round(ms, 4), "gens_per_s": round(1000.0 / ms, 2), "generations_timed": generations,
            "reference_ms_per_gen": 5.1406, "reference_hardware": "RTX PRO 6000 Blackwell",
            "speedup_vs_reference": round(5.1406 / ms, 3)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-dtlz1", action="store_true")
    ap.add_argument("--batched-instances", type=int, default=256)
    ap.add_argument("--batched-generations", type=int, default=300)
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-rvea", action="store_true")
    ap.add_argument("--rvea-generations", type=int, default=5)
    ap.add_argument("--serve-seconds", type=float, default=120.0)
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-plain-sample", type=int, default=24)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: each population runs on a single GPU and there is "
              "no multi-GPU path; use --gpus 1", file=sys.stderr)
        return 2
    import torch
    import cuda_moea as cm
    from cuda_moea import _lib

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    K, W = args.steps, args.warmup
    stream = torch.cuda.current_stream(dev)

    def config1_run():
        a = cm.NSGA3(population_size=N, max_generations=W + K + 1,
                     problem=cm.QuadraticCenters(D, M, seed=0), seed=0, print_progress=False)
        a.initialize()
        if W:
            a._call(_lib.lib.moea_run_step, W)
        stream.synchronize()
        return a

    alg = config1_run()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    alg._call(_lib.lib.moea_run_step, K)
    t1.record(stream)
    t1.synchronize()
    ms = t0.elapsed_time(t1)
    ms_per_gen = ms / K
    max_front = int(alg._stats[0].item())
    alg.synchronize()                                    # raises if a peel launch failed
    del alg

    # pass 2: the same generations W .. W+K-1 of an identical run (same seeds,
    # bit-identical data) with HIP events around every stage on the run stream
    alg = config1_run()
    _lib.profile_enable(True)
    alg._call(_lib.lib.moea_run_step, K)
    stream.synchronize()
    stages = _lib.profile_read()
    _lib.profile_enable(False)
    st = alg._stats.cpu().numpy()
    n_assigned = int(st[1]) + int(st[2])                 # rows in fronts <= max_front (last generation)
    stage_ms = {k: round(v[0] / max(v[1], 1), 4) for k, v in stages.items() if v[1]}
    nm = 2 * N
    words = (nm + 63) // 64

    def avg_s(name):
        t, n = stages[name]
        return t / max(n, 1) / 1e3

    # nds_dominance: every unordered pair once, M lane instructions each (M-1
    # compares into a lane mask, one bit shift-in); bytes: the block-lower-
    # triangular bitmask written plus the sorted objectives read
    dom_avg_s = avg_s("nds_dominance")
    lane_ops = nm * (nm - 1) // 2 * M
    dom_bytes = 64 * 8 * words * (words + 1) // 2 + nm * M * 4
    dominance = {"bound": "valu", "kernel": "nds_dominance",
                 "achieved": round(lane_ops / dom_avg_s / 1e9, 1), "peak": round(VALU_PEAK_CMP / 1e9, 1),
                 "unit": "Glane-ops/s", "frac": round(lane_ops / dom_avg_s / VALU_PEAK_CMP, 4),
                 "work_basis": "unordered pairs x M VALU lane-ops (the kernel's own instruction count)",
                 "work_per_launch": lane_ops, "avg_launch_ms": round(dom_avg_s * 1e3, 4),
                 "frac_reference_work": round(M * nm * (nm - 1) / dom_avg_s / VALU_PEAK_CMP, 4),
                 "reference_work_basis": "M x N_mix x (N_mix - 1) compares at one compare per lane-cycle "
                                         "(can exceed 1: the sort order settles half of them unexecuted)",
                 "algorithmic_bytes": dom_bytes, "traffic": _traffic_from_profiles("nds_dominance"),
                 "hbm_frac": round(dom_bytes / dom_avg_s / HBM_PEAK, 4)}
    # nds_peel: each row assigned a front must be checked through every word of
    # its dominator row up to its block's tie bound (about W/2 words for rows
    # spread over the sort order): n_assigned x (W + 1) / 2 x 8 B read once
    peel_avg_s = avg_s("nds_peel")
    peel_bytes = n_assigned * (words + 1) // 2 * 8
    peel = {"bound": "hbm", "kernel": "nds_peel",
            "achieved": round(peel_bytes / peel_avg_s / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(peel_bytes / peel_avg_s / HBM_PEAK, 4),
            "traffic": _traffic_from_profiles("nds_peel"),
            "algorithmic_bytes": peel_bytes, "avg_launch_ms": round(peel_avg_s * 1e3, 4),
            "work_basis": f"{n_assigned} rows in fronts up to the last (last generation) x (W+1)/2 words x 8 B; "
                          "latency-bound: one grid barrier and one or two dependent HBM reads per front"}
    roofline = peel if peel_avg_s > dom_avg_s else dominance
    other = dominance if roofline is peel else peel

    batched = None
    if not args.no_batched:
        batched = batched_config(args.batched_generations, args.batched_instances, args.plain_sample)

    selection = selection_bench(20)

    rvea = None
    if not args.no_rvea:
        rvea = rvea_config(args.rvea_generations, 2)

    serving = None
    if not args.no_serving:
        serving = serving_config(args.serve_seconds, args.serve_rate, args.serve_plain_sample,
                                 int(args.serve_budget_gb * 1e9))

    dtlz1 = None
    if not args.no_dtlz1:
        dtlz1 = dtlz1_comparison(50, 5)

    cpu = None
    config0 = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_threads)
        config0 = config0_cpu(args.cpu_threads)

    line = {
        "metric": "gens/sec NSGA-III N=32768 M=3; batched runs/sec + p99 latency vs plain",
        "value": round(1000.0 / ms_per_gen, 3),
        "unit": "gens/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_per_gen, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: quadratic-center objectives, centers and population seeded (seed 0)",
        "config": {"workload": "NSGA-III single population, quadratic-center objectives",
                   "N": N, "M": M, "D": D, "K_directions": alg.k, "generations_timed": f"{W}-{W + K - 1}",
                   "last_front_index": max_front},
        "roofline": roofline,
        "roofline_other": other,
        "stage_ms_per_generation": stage_ms,
        "stage_window": f"generations {W}-{W + K - 1} (the timed ones, second identical run with stage events)",
        "batched": batched,
        "selection_bench": selection,
        "serving": serving,
        "rvea": rvea,
        "reference_comparison": dtlz1,
        "cpu_baseline": cpu,
        "config0_cpu": config0,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
