"""Batched NSGA-III: many independent runs advanced together.

Each run is given as an uninitialized ``cuda_moea.NSGA3`` with a
``QuadraticCenters`` problem (tournament mating, SBX, polynomial mutation, the
defaults).  The runs may differ in population size, objective count,
dimension, seed and operator parameters, but share the number of generations.
libmoea_hip's ``moea_batch_*`` advances all of them with one launch sequence
per generation, and each run's population is bit-identical to running that
``NSGA3`` alone.  The reference has no batched engine (each
EvolutionaryAlgorithm runs one population, NS/algorithm/algorithm.cu:104-329).
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
from .problem import QuadraticCenters


class BatchNSGA3:
    def __init__(self, runs: Sequence[NSGA3]) -> None:
        runs = list(runs)
        if not runs:
            raise ValueError("BatchNSGA3 needs at least one run")
        for r in runs:
            if not isinstance(r, NSGA3):
                raise TypeError("BatchNSGA3 takes cuda_moea.NSGA3 runs")
            if not isinstance(r.problem, QuadraticCenters):
                raise NotImplementedError("BatchNSGA3 supports QuadraticCenters problems")
            if not isinstance(r.mating, TournamentMating) or not isinstance(r.mutation, PolynomialMutation):
                raise NotImplementedError("BatchNSGA3 runs use tournament mating and polynomial mutation")
            if r.initial_population is not None:
                raise NotImplementedError("BatchNSGA3 draws its initial populations")
            if r.population_size <= 0 or r.population_size % 2:
                raise ValueError("The current NSGA-III backend requires an even population size")
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
            p = r.problem
            n, d, m = r.population_size, int(p.dimension), int(p.objectives)
            if m > _lib.MAX_OBJECTIVES:
                raise ValueError(f"this engine supports at most {_lib.MAX_OBJECTIVES} objectives")
            bounds = np.empty(2 * d, np.float32)
            bounds[0::2] = np.asarray(p.lower_bounds, np.float32)
            bounds[1::2] = np.asarray(p.upper_bounds, np.float32)
            v = r._directions_host(n, m)
            centers = p.centers.to(dev).contiguous()
            tb = torch.from_numpy(bounds).to(dev)
            tv = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
            self._keep += [centers, tb, tv]
            e = arr[i]
            e.n, e.m, e.d, e.k, e.max_generations = n, m, d, v.shape[1], r.max_generations
            e.seed = r._run_seed()
            e.centers, e.bounds, e.directions = centers.data_ptr(), tb.data_ptr(), tv.data_ptr()
            c, mu = r.crossover, r.mutation
            e.sbx_eta[0], e.sbx_eta[1] = c.eta_initial, c.eta_final
            e.sbx_prob, e.sbx_copy_prob = c.probability, c.variable_copy_probability
            e.pm_eta[0], e.pm_eta[1] = mu.eta_initial, mu.eta_final
            e.pm_prob = mu.probability
        self._arr = arr
        nbytes = lib.moea_batch_workspace_bytes(arr, count)
        if nbytes == 0:
            check(-1)
        self._ws = K.workspace(nbytes, dev)
        stream = torch.cuda.current_stream(dev).cuda_stream
        check(lib.moea_batch_create(arr, count, self._ws.data_ptr(), self._ws.numel(), stream,
                                    ctypes.byref(self._handle)))
        check(lib.moea_batch_init(self._handle))
        self._initialized = True

    def step(self, generations: int = 1) -> None:
        if not self._initialized:
            self.initialize()
        check(lib.moea_batch_step(self._handle, int(generations)))

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

    def _view(self, ptr: int, count: int, dtype: torch.dtype) -> torch.Tensor:
        off = ptr - self._ws.data_ptr()
        nbytes = count * torch.tensor([], dtype=dtype).element_size()
        return self._ws[off:off + nbytes].view(dtype)

    def result(self, index: int, *, copy: bool = True) -> Result:
        if not self._initialized:
            raise RuntimeError("BatchNSGA3 has not been initialized")
        pop = _lib.PopulationC()
        fronts = ctypes.c_void_p()
        check(lib.moea_batch_view(self._handle, int(index), ctypes.byref(pop), ctypes.byref(fronts)))
        r = self.runs[index]
        n, d, m = r.population_size, int(r.problem.dimension), int(r.problem.objectives)
        torch.cuda.current_stream(self.device).synchronize()
        x = self._view(pop.x, n * d, torch.float32).view(n, d)
        f = self._view(pop.f, m * n, torch.float32).view(m, n).t()
        cv = self._view(pop.cv, n, torch.float32)
        fr = self._view(fronts.value, n, torch.int32)
        if copy:
            x, f, cv, fr = x.clone(), f.contiguous(), cv.clone(), fr.clone()
        return Result(x, f, cv, fr, n, self._total_ms)

    def __del__(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            lib.moea_batch_destroy(h)
            self._handle = ctypes.c_void_p()
