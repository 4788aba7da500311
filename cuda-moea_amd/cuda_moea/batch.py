"""Batched NSGA-III: many independent runs advanced together.

Each run is given as an uninitialized ``cuda_moea.NSGA3`` with a
``QuadraticCenters`` or ``SeededMLP`` problem (tournament mating, SBX,
polynomial mutation, the defaults).  The runs may differ in population size,
objective count, dimension, problem, seed and operator parameters, but share
the number of generations.  libmoea_hip's ``moea_batch_*`` advances all of
them with one launch sequence per generation, and each run's population is
bit-identical to running that ``NSGA3`` alone.  The reference has no batched
engine (each EvolutionaryAlgorithm runs one population,
NS/algorithm/algorithm.cu:104-329).
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from ._specs import PolynomialMutation, TournamentMating
from . import kernels as K
from .algorithm import NSGA3, Result
from .problem import QuadraticCenters, SeededMLP


def check_run(r) -> None:
    """Raise unless ``r`` is a run the batched engine can advance."""
    if not isinstance(r, NSGA3):
        raise TypeError("batched runs are cuda_moea.NSGA3 instances")
    if not isinstance(r.problem, (QuadraticCenters, SeededMLP)):
        raise NotImplementedError("batched runs support QuadraticCenters and SeededMLP problems")
    if not isinstance(r.mating, TournamentMating) or not isinstance(r.mutation, PolynomialMutation):
        raise NotImplementedError("batched runs use tournament mating and polynomial mutation")
    if r.initial_population is not None:
        raise NotImplementedError("batched runs draw their initial populations")
    if r.population_size <= 0 or r.population_size % 2:
        raise ValueError("The current NSGA-III backend requires an even population size")
    if int(r.problem.objectives) > _lib.MAX_OBJECTIVES:
        raise ValueError(f"this engine supports at most {_lib.MAX_OBJECTIVES} objectives")


def fill_problem(e: "_lib.BatchProblem", r, dev: torch.device, keep: list) -> None:
    """Fill one moea_batch_problem from an NSGA3 run; the device tensors it
    points at are appended to ``keep`` (they must outlive the run)."""
    p = r.problem
    n, d, m = r.population_size, int(p.dimension), int(p.objectives)
    bounds = np.empty(2 * d, np.float32)
    bounds[0::2] = np.asarray(p.lower_bounds, np.float32)
    bounds[1::2] = np.asarray(p.upper_bounds, np.float32)
    v = r._directions_host(n, m)
    tb = torch.from_numpy(bounds).to(dev)
    tv = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    keep += [tb, tv]
    e.n, e.m, e.d, e.k, e.max_generations = n, m, d, v.shape[1], r.max_generations
    e.seed = r._run_seed()
    e.bounds, e.directions = tb.data_ptr(), tv.data_ptr()
    if isinstance(p, SeededMLP):
        w1 = p.w1.to(dev).contiguous()
        w2 = p.w2.to(dev).contiguous()
        keep += [w1, w2]
        e.problem, e.w1, e.w2, e.hidden = _lib.PROBLEM_MLP, w1.data_ptr(), w2.data_ptr(), p.hidden
    else:
        centers = p.centers.to(dev).contiguous()
        keep.append(centers)
        e.problem, e.centers = _lib.PROBLEM_QUADRATIC, centers.data_ptr()
    c, mu = r.crossover, r.mutation
    e.sbx_eta[0], e.sbx_eta[1] = c.eta_initial, c.eta_final
    e.sbx_prob, e.sbx_copy_prob = c.probability, c.variable_copy_probability
    e.pm_eta[0], e.pm_eta[1] = mu.eta_initial, mu.eta_final
    e.pm_prob = mu.probability


def view_result(ws: torch.Tensor, pop, fronts: int, r, total_ms: float, copy: bool) -> Result:
    """Result of run ``r`` from device views into the workspace ``ws``."""
    def at(ptr: int, count: int, dtype: torch.dtype) -> torch.Tensor:
        off = ptr - ws.data_ptr()
        nbytes = count * torch.tensor([], dtype=dtype).element_size()
        return ws[off:off + nbytes].view(dtype)

    n, d, m = r.population_size, int(r.problem.dimension), int(r.problem.objectives)
    x = at(pop.x, n * d, torch.float32).view(n, d)
    f = at(pop.f, m * n, torch.float32).view(m, n).t()
    cv = at(pop.cv, n, torch.float32)
    fr = at(fronts, n, torch.int32)
    if copy:
        x, f, cv, fr = x.clone(), f.contiguous(), cv.clone(), fr.clone()
    return Result(x, f, cv, fr, n, total_ms)


class BatchNSGA3:
    def __init__(self, runs: Sequence[NSGA3], *, niching: tuple[int, int] = (0, 0), graph: bool = False) -> None:
        """``niching`` = (mode, lds_cap) of moea_niching: the niching path of
        every selection of the batch ((0, 0): automatic).  ``graph``: replay
        each generation from a captured HIP graph (moea_batch_set_graph)."""
        runs = list(runs)
        self._niching = (int(niching[0]), int(niching[1]))
        self._graph = bool(graph)
        if not runs:
            raise ValueError("BatchNSGA3 needs at least one run")
        for r in runs:
            check_run(r)
            if r.max_generations != runs[0].max_generations:
                raise ValueError("every run of a batch needs the same max_generations")
            if r.cuda.device_id != runs[0].cuda.device_id:
                raise ValueError("every run of a batch must use the same device")
        self.runs = runs
        self.max_generations = runs[0].max_generations
        self.device = torch.device("cuda", runs[0].cuda.device_id)
        self._handle = ctypes.c_void_p()
        self._initialized = False
        self._total_ms = 0.0

    def initialize(self) -> None:
        if self._initialized:
            return
        if not torch.cuda.is_available():
            raise RuntimeError("cuda_moea requires a ROCm GPU")
        dev = self.device
        count = len(self.runs)
        arr = (_lib.BatchProblem * count)()
        self._keep = []
        for i, r in enumerate(self.runs):
            fill_problem(arr[i], r, dev, self._keep)
            arr[i].niching.mode, arr[i].niching.lds_cap = self._niching
        self._arr = arr
        nbytes = lib.moea_batch_workspace_bytes(arr, count)
        if nbytes == 0:
            check(-1)
        self._ws = K.workspace(nbytes, dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(lib.moea_batch_create(arr, count, self._ws.data_ptr(), self._ws.numel(), stream,
                                    ctypes.byref(self._handle)))
        check(lib.moea_batch_init(self._handle))
        check(lib.moea_batch_set_graph(self._handle, 1 if self._graph else 0))
        self._initialized = True

    def step(self, generations: int = 1) -> None:
        if not self._initialized:
            self.initialize()
        check(lib.moea_batch_step(self._handle, int(generations)))

    def autotune(self, trials: int = 3) -> dict:
        """Choose the front-peeling plan by timing this batch's own next
        generations under each candidate (moea_batch_autotune); the results
        equal an untuned run.  Returns the chosen (capacity divisor, rows per
        workgroup), the ms per generation of every candidate and the
        generations used."""
        if not self._initialized:
            self.initialize()
        chosen = ctypes.c_int(-1)
        ms = (ctypes.c_double * 8)()
        ran = lib.moea_batch_autotune(self._handle, int(trials), ctypes.byref(chosen), ms, 8)
        if ran < 0:
            check(ran)
        div, rows = ctypes.c_int(), ctypes.c_int()
        check(lib.moea_batch_peel_plan(self._handle, ctypes.byref(div), ctypes.byref(rows)))
        return {"chosen": chosen.value, "plan": (div.value, rows.value), "ms_per_generation": list(ms)[:4],
                "generations": ran}

    def graph_info(self) -> dict:
        """Nodes of the captured generation graph (zeros before a capture)."""
        k, ms, caps = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        n = lib.moea_batch_graph_info(self._handle, ctypes.byref(k), ctypes.byref(ms), ctypes.byref(caps))
        if n < 0:
            check(n)
        return {"nodes": n, "kernel_nodes": k.value, "memset_nodes": ms.value, "captures": caps.value}

    @property
    def generation(self) -> int:
        return lib.moea_batch_generation(self._handle) if self._initialized else 0

    @property
    def finished(self) -> bool:
        return self._initialized and self.generation >= self.max_generations

    def run(self, *, copy: bool = True) -> list[Result]:
        if not self._initialized:
            self.initialize()
        stream = torch.cuda.current_stream(self.device)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        self.step(self.max_generations - self.generation)
        stop.record(stream)
        stop.synchronize()
        self._total_ms = float(start.elapsed_time(stop))
        return [self.result(i, copy=copy) for i in range(len(self.runs))]

    def result(self, index: int, *, copy: bool = True) -> Result:
        if not self._initialized:
            raise RuntimeError("BatchNSGA3 has not been initialized")
        pop = _lib.PopulationC()
        fronts = ctypes.c_void_p()
        check(lib.moea_batch_view(self._handle, int(index), ctypes.byref(pop), ctypes.byref(fronts)))
        # waits for the queued generations and raises if one of them failed on the device
        check(lib.moea_batch_check(self._handle))
        return view_result(self._ws, pop, fronts.value, self.runs[index], self._total_ms, copy)

    def __del__(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            lib.moea_batch_destroy(h)
            self._handle = ctypes.c_void_p()
