"""Python facade with the reference's API (python/cuda_moea/algorithm.py:16-165).

The generation loop itself is native: ``moea_run_init`` / ``moea_run_step`` in
libmoea_hip.so run EvolutionaryAlgorithm::initialize / step
(NS/algorithm/algorithm.cu:104-286) as stream-ordered launches with no host
synchronization.  Python owns the device buffers (torch tensors) and the
configuration.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np
import torch

from . import _lib
from ._lib import MoeaError, check, lib
from ._specs import (
    AdaptiveRVEADirections, CudaConfig, DasDennisDirections, NoMutation,
    NSGA3EnvironmentSelector, PolynomialMutation, PythonCrossover, PythonMating, PythonMutation,
    PythonEnvironmentSelector, PythonProblem, PythonReferenceDirections, RandomMating, RVEAEnvironmentSelector, SBX,
    TournamentMating, UserDefinedDirections,
)
from . import kernels as K
from .native import EVALUATE as _NATIVE_EVALUATE, PREPARE_PARENT as _NATIVE_PREPARE_PARENT, NativeProblem
from .problem import DTLZ2, DTLZProblem, QuadraticCenters, SeededMLP
from .record import SnapshotRecorder


@dataclass(frozen=True)
class Population:
    variables: Any
    objectives: Any
    constraints: Any


@dataclass(frozen=True)
class Result(Population):
    auxiliary_indices: Any
    active_count: int
    total_ms: float


def _checked(value: Any, name: str, device: int, dtype: torch.dtype) -> torch.Tensor:
    # adapters.cpp:101-112: a tensor on the run's device, converted to dtype
    if not isinstance(value, torch.Tensor):
        raise ValueError(f"{name} must be a torch.Tensor")
    if not value.is_cuda or value.device.index != device:
        raise ValueError(f"{name} must be on the algorithm CUDA device")
    return value.to(dtype).contiguous()


def _cv_quant_valid(sel: Any) -> bool:
    # validate_cv_quant_params (cv_quantization.cu:93-112)
    clip, alpha, eps = float(sel.cv_clip_upper), float(sel.cv_log_alpha), float(sel.feasibility_epsilon)
    if not (clip > 0.0 and math.isfinite(clip)) or not (alpha >= 0.0 and math.isfinite(alpha)):
        return False
    if not (eps >= 0.0 and math.isfinite(eps)):
        return False
    return alpha == 0.0 or (math.log1p(alpha * clip) > 0.0 and math.isfinite(math.log1p(alpha * clip)))


def _unsupported(obj: Any, what: str) -> NotImplementedError:
    return NotImplementedError(f"{type(obj).__name__} is not supported as {what} by this engine")


class Algorithm:
    def __init__(self, algorithm: str = "NSGA3", *, population_size: int = 1024,
                 max_generations: int = 100, problem: Any | None = None,
                 mating: Any | None = None, crossover: Any | None = None,
                 mutation: Any | None = None, reference_directions: Any | None = None,
                 environment_selector: Any | None = None, initial_population: Any | None = None,
                 cuda: CudaConfig | None = None, device: int | str | None = None,
                 seed: int | None = None, enable_warmup: bool | None = None,
                 progress_interval: int = 100, print_progress: bool = True,
                 save_data: str | Path | None = None, save_interval: int = 1) -> None:
        kind = algorithm.upper().replace("-", "")
        if kind not in {"NSGA3", "RVEA"}:
            raise ValueError("algorithm must be 'NSGA3' or 'RVEA'")
        cuda = CudaConfig() if cuda is None else cuda
        if device is not None:
            cuda.device_id = _device_index(device)
        if seed is not None:
            cuda.seed = seed
        if enable_warmup is not None:
            cuda.enable_warmup = enable_warmup
        if int(save_interval) <= 0:
            # algorithm_factory.cu:84-87 checks the interval whether or not saving is on
            raise ValueError("saveData generation interval must be positive")
        # python/cuda_moea/algorithm.py:72-74: any non-None save_data enables saving
        self._recorder = (SnapshotRecorder(str(save_data or "cuda_moea_data"), int(save_interval))
                          if save_data is not None else None)
        self._front0: np.ndarray | None = None
        self.kind = kind
        self.cuda = cuda
        self.population_size = int(population_size)
        self.max_generations = int(max_generations)
        self.problem = DTLZ2() if problem is None else problem
        self.mating = (TournamentMating() if kind == "NSGA3" else RandomMating()) if mating is None else mating
        self.crossover = SBX() if crossover is None else crossover
        self.mutation = PolynomialMutation() if mutation is None else mutation
        self.reference_directions = ((DasDennisDirections() if kind == "NSGA3" else AdaptiveRVEADirections())
                                     if reference_directions is None else reference_directions)
        self.environment_selector = ((NSGA3EnvironmentSelector() if kind == "NSGA3" else RVEAEnvironmentSelector())
                                     if environment_selector is None else environment_selector)
        self.initial_population = initial_population
        self.progress_interval = progress_interval
        self.print_progress = print_progress
        self._validate_strategies()
        # a native problem is created now, as the reference's algorithm
        # constructor does, so that its factory's checks raise here
        self._native = self.problem._instantiate() if isinstance(self.problem, NativeProblem) else None
        self._initialized = False
        self._total_ms = 0.0
        self._resets = 0
        self._error: BaseException | None = None

    # ------------------------------------------------------------ config ----
    def _validate_strategies(self) -> None:
        if not isinstance(self.mating, (TournamentMating, RandomMating, PythonMating)):
            raise _unsupported(self.mating, "mating")
        if isinstance(self.crossover, SBX):
            self.crossover.validate()
        elif not isinstance(self.crossover, PythonCrossover):
            raise _unsupported(self.crossover, "crossover")
        if isinstance(self.mutation, PolynomialMutation):
            self.mutation.validate()
        elif not isinstance(self.mutation, (NoMutation, PythonMutation)):
            raise _unsupported(self.mutation, "mutation")
        rd = self.reference_directions
        if not isinstance(rd, (DasDennisDirections, AdaptiveRVEADirections, UserDefinedDirections,
                               PythonReferenceDirections)):
            raise _unsupported(rd, "reference directions")
        if isinstance(rd, DasDennisDirections) and rd.partitions < 0:
            raise ValueError("Das-Dennis partitions cannot be negative")
        sel = self.environment_selector
        if isinstance(sel, PythonEnvironmentSelector):
            if self.kind == "NSGA3" and isinstance(rd, AdaptiveRVEADirections):
                raise ValueError("NSGA-III takes Das-Dennis or user-defined directions")
        elif self.kind == "NSGA3":
            if not isinstance(sel, NSGA3EnvironmentSelector):
                raise _unsupported(sel, "NSGA-III environment selector")
            if not sel.sparse_ratio > 0.0:
                raise ValueError("NSGA-III sparse_ratio must be positive")
            if sel.cv_bins > 0 and not _cv_quant_valid(sel):
                # nsga3_environment_selector.cu:216-220
                raise ValueError("Invalid NSGA-III CV quantization configuration")
            if isinstance(rd, AdaptiveRVEADirections):
                raise ValueError("NSGA-III takes Das-Dennis or user-defined directions")
        else:
            if not isinstance(sel, RVEAEnvironmentSelector):
                raise _unsupported(sel, "RVEA environment selector")
            if not sel.alpha > 0.0:
                raise ValueError("RVEA alpha must be positive")
        if not isinstance(self.problem, (QuadraticCenters, SeededMLP, PythonProblem, NativeProblem)):
            raise _unsupported(self.problem, "problem")

    # ------------------------------------------------------------ device ----
    def initialize(self) -> None:
        if self._initialized:
            return
        if self.population_size <= 0 or self.max_generations <= 0:
            raise ValueError("Algorithm requires positive population_size and max_generations")
        if not torch.cuda.is_available():
            raise RuntimeError("cuda_moea requires a ROCm GPU")
        p = self.problem
        n, d, m = self.population_size, int(p.dimension), int(p.objectives)
        if d <= 0 or m <= 0 or len(p.lower_bounds) != d or len(p.upper_bounds) != d:
            raise ValueError("Problem must provide positive dimensions and one lower/upper bound per decision variable")
        if m > _lib.MAX_OBJECTIVES:
            raise ValueError(f"this engine supports at most {_lib.MAX_OBJECTIVES} objectives")
        if self.kind == "NSGA3" and n % 2 != 0:
            raise ValueError("The current NSGA-III backend requires an even population size")
        dev = torch.device("cuda", self.cuda.device_id)
        self.device = dev
        self.n, self.d, self.m = n, d, m
        lb = np.asarray(p.lower_bounds, np.float32)
        ub = np.asarray(p.upper_bounds, np.float32)
        bounds = np.empty(2 * d, np.float32)
        bounds[0::2], bounds[1::2] = lb, ub
        self._bounds = torch.from_numpy(bounds).to(dev)
        self._lb, self._ub = lb, ub
        v = self._directions_host(n, m)
        k = v.shape[1]
        self.k = k
        self._directions = torch.from_numpy(v).to(dev)
        self._directions0 = self._directions.clone()
        self._pops = [K.Population.empty(n, d, m, dev) for _ in range(3)]
        self._fronts = [torch.full((n,), -1, dtype=torch.int32, device=dev) for _ in range(2)]
        self._parent_idx = torch.empty(n, dtype=torch.int32, device=dev)
        self._ideal = torch.empty(m, dtype=torch.float32, device=dev)
        self._gamma = torch.empty(k, dtype=torch.float32, device=dev)
        self._active = torch.full((1,), n, dtype=torch.int32, device=dev)
        self._scratch = torch.empty(max(m, 16), dtype=torch.float32, device=dev)
        self._stats = torch.zeros(8, dtype=torch.int32, device=dev)
        self._stats_init = torch.zeros(8, dtype=torch.int32, device=dev)
        alg = _lib.ALGORITHM_NSGA3 if self.kind == "NSGA3" else _lib.ALGORITHM_RVEA
        self._ws = K.workspace(lib.moea_run_workspace_bytes(alg, n, m, k), dev)
        a = _lib.RunArgs()
        a.algorithm, a.n, a.m, a.d, a.k, a.max_generations = alg, n, m, d, k, self.max_generations
        for i, pop in enumerate(self._pops):
            a.pops[i] = pop.c()
        a.current = 0
        a.fronts[0], a.fronts[1] = self._fronts[0].data_ptr(), self._fronts[1].data_ptr()
        a.fronts_current = 0
        a.parent_indices = self._parent_idx.data_ptr()
        a.bounds = self._bounds.data_ptr()
        a.directions, a.directions0 = self._directions.data_ptr(), self._directions0.data_ptr()
        a.gamma, a.ideal = self._gamma.data_ptr(), self._ideal.data_ptr()
        a.active_device, a.adapt_scratch = self._active.data_ptr(), self._scratch.data_ptr()
        a.stats, a.stats_init = self._stats.data_ptr(), self._stats_init.data_ptr()
        a.mating = _lib.MATING_TOURNAMENT if isinstance(self.mating, TournamentMating) else _lib.MATING_RANDOM
        if isinstance(self.mutation, PolynomialMutation):
            a.mutation = _lib.MUTATION_POLYNOMIAL
            a.pm_eta[0], a.pm_eta[1] = self.mutation.eta_initial, self.mutation.eta_final
            a.pm_prob = self.mutation.probability
        else:
            a.mutation = _lib.MUTATION_NONE
        c = self.crossover
        if isinstance(c, SBX):
            a.sbx_eta[0], a.sbx_eta[1] = c.eta_initial, c.eta_final
            a.sbx_prob, a.sbx_copy_prob = c.probability, c.variable_copy_probability
        if self.kind == "RVEA":
            sel = self.environment_selector
            # the built-in selector's penalty exponent; a Python selector computes its own
            a.rvea_alpha = sel.alpha if isinstance(sel, RVEAEnvironmentSelector) else 2.0
            freq = getattr(self.reference_directions, "frequency", 1.0)
            if isinstance(self.reference_directions, AdaptiveRVEADirections) and not 0.0 < freq <= 1.0:
                raise ValueError("RVEA reference adaptation frequency must be in (0, 1]")
            # fixed (user-defined) directions never adapt: an interval past the run
            a.rvea_frequency = freq if isinstance(self.reference_directions, AdaptiveRVEADirections) else 1e9
        a.seed = self._run_seed()
        sel = self.environment_selector
        if isinstance(sel, NSGA3EnvironmentSelector) and sel.cv_bins > 0:
            a.cv_quant = K.cv_quantization(sel.cv_bins, sel.cv_clip_upper, sel.cv_log_alpha, sel.feasibility_epsilon)
        self._by_ptr = {pop.x.data_ptr(): pop for pop in self._pops}
        self._setup_problem(a)
        self._setup_hooks(a)
        if self.initial_population is not None:
            init = torch.as_tensor(self.initial_population, dtype=torch.float32)
            if tuple(init.shape) != (n, d) and init.numel() != n * d:
                raise ValueError("initial_population must have shape (population_size, dimension)")
            init = init.reshape(n, d)
            if (not torch.isfinite(init).all() or (init < torch.from_numpy(lb)).any()
                    or (init > torch.from_numpy(ub)).any()):
                raise ValueError("initial_population contains a non-finite or out-of-bounds value")
            self._pops[0].x.copy_(init.to(dev))
            a.initial_x_given = 1
        a.workspace, a.workspace_bytes = self._ws.data_ptr(), self._ws.numel()
        a.stream = torch.cuda.current_stream(dev).cuda_stream
        if self._native is not None:
            self._native.initialize(a.stream, a.seed, self._bounds.data_ptr())
        self._args = a
        self._call(lib.moea_run_init)
        if self._python_selector:
            self._selector_start()
        self._initialized = True
        self._total_ms = 0.0
        self._front0 = None
        if self._recorder is not None:
            # algorithm.cu:208-214: metadata, bounds and initial directions, then generation 0
            self._recorder.initialize(
                algorithm_name="NSGA-III" if self.kind == "NSGA3" else "RVEA",
                problem_name=getattr(p, "name", type(p).__name__), population_size=n, dimension=d,
                objective_count=m, constraint_count=int(getattr(p, "constraints", 0)),
                max_generations=self.max_generations, random_seed=int(self.cuda.seed),
                lower_bounds=lb, upper_bounds=ub, initial_directions=self._directions0.cpu().numpy())
            self._snapshot(0)

    def _run_seed(self) -> int:
        s = int(self.cuda.seed) & (2**64 - 1)
        return (s + self._resets * 0x9E3779B97F4A7C15) & (2**64 - 1)

    def _directions_host(self, n: int, m: int) -> np.ndarray:
        rd = self.reference_directions
        if isinstance(rd, PythonReferenceDirections):
            # adapters.cpp:294-301, 325-335: (count, objective_count) on the device
            out = _checked(rd.initialize(n, m, {"device": self.cuda.device_id}), "reference directions",
                           self.cuda.device_id, torch.float32)
            if out.dim() != 2 or out.shape[1] != m or out.shape[0] == 0:
                raise ValueError("reference directions must have shape (count, objective_count)")
            return np.ascontiguousarray(out.t().cpu().numpy())
        if isinstance(rd, UserDefinedDirections):
            # adapters.cpp:557-562: a (count, objective_count) tensor, stored
            # objective-major; reference_directions.cu:189-206 checks the sizes
            vals = rd.values.detach().cpu() if isinstance(rd.values, torch.Tensor) else rd.values
            vals = np.asarray(vals, dtype=np.float32)
            if vals.ndim != 2:
                raise ValueError("user directions must be a 2-D tensor")
            if vals.size == 0:
                raise ValueError("User directions must have shape (objective_count, count)")
            if vals.shape[1] != m:
                raise ValueError("User direction objective count does not match the problem")
            vals = np.array(vals.T, dtype=np.float32, copy=True)
            if rd.normalize:
                for j in range(vals.shape[1]):
                    ss = np.float32(0.0)
                    for c in range(m):
                        ss = np.float32(ss + vals[c, j] * vals[c, j])
                    if ss <= 0:
                        raise ValueError("Reference directions must be non-zero")
                    inv = np.float32(1.0) / np.sqrt(ss)
                    vals[:, j] = vals[:, j] * inv
            return np.ascontiguousarray(vals)
        if isinstance(rd, AdaptiveRVEADirections):
            return K.das_dennis(n, m, 0, kind=2)
        return K.das_dennis(n, m, rd.partitions)

    def _setup_problem(self, a: _lib.RunArgs) -> None:
        p = self.problem
        pd = _lib.ProblemDesc()
        if isinstance(p, QuadraticCenters):
            self._centers = p.centers.to(self.device).contiguous()
            pd.kind, pd.centers = _lib.PROBLEM_QUADRATIC, self._centers.data_ptr()
        elif isinstance(p, SeededMLP):
            self._w1 = p.w1.to(self.device).contiguous()
            self._w2 = p.w2.to(self.device).contiguous()
            self._hwork = torch.empty(self.n, p.hidden, dtype=torch.float32, device=self.device)
            pd.kind, pd.w1, pd.w2 = _lib.PROBLEM_MLP, self._w1.data_ptr(), self._w2.data_ptr()
            pd.hidden, pd.hwork = p.hidden, self._hwork.data_ptr()
        elif isinstance(p, NativeProblem):
            # the plugin's evaluate and prepare_parent are called by the run
            # loop through the loader's C entry points
            pd.kind, pd.callback, pd.user = _lib.PROBLEM_CALLBACK, _NATIVE_EVALUATE, self._native.handle
            pd.prepare_parent = _NATIVE_PREPARE_PARENT
        elif isinstance(p, DTLZProblem) and p._kind in _lib.DTLZ_KINDS and 2 <= self.m <= self.d:
            # built-in objective kernels, as the reference's DTLZ problems are native
            pd.kind = _lib.DTLZ_KINDS[p._kind]
            pd.cv_activation_ratio = p.constraint_activation_ratio
        else:
            if hasattr(p, "initialize"):
                p.initialize({"device": self.cuda.device_id})
            self._callback = _lib.EVALUATE_FN(self._evaluate_python)
            pd.kind, pd.callback = _lib.PROBLEM_CALLBACK, self._callback
            if hasattr(p, "prepare_parent"):
                self._prepare_cb = _lib.PREPARE_PARENT_FN(self._prepare_parent_python)
                pd.prepare_parent = self._prepare_cb
        a.problem = pd

    def _context(self, generation: int, max_generations: int) -> dict[str, Any]:
        return {"generation": generation, "max_generations": max_generations, "device": self.cuda.device_id}

    def _pop_dict(self, pop_ptr) -> dict[str, Any]:
        return self._pop_view(self._by_ptr[pop_ptr.contents.x])

    def _pop_view(self, pop: K.Population) -> dict[str, Any]:
        # adapters.cpp:80-88: objectives as an (n, m) view of the (m, n) storage
        return {"variables": pop.x, "objectives": pop.f.t(), "constraints": pop.cv,
                "bounds": self._bounds.view(self.d, 2)}

    def _setup_hooks(self, a: _lib.RunArgs) -> None:
        """Python strategies as native hooks of the generation loop
        (adapters.cpp:196-338); built-in strategies stay on the GPU."""
        info = {"name": "NSGA-III" if self.kind == "NSGA3" else "RVEA", "population_size": self.n,
                "dimension": self.d, "objective_count": self.m, "max_generations": self.max_generations,
                "device": self.cuda.device_id}
        h = _lib.StrategyHooks()
        used = False
        if isinstance(self.mating, PythonMating):
            h.mate, used = _lib.MATE_FN(self._mate_python), True
        if isinstance(self.crossover, PythonCrossover):
            h.crossover, used = _lib.CROSSOVER_FN(self._crossover_python), True
        if isinstance(self.mutation, PythonMutation):
            h.mutate, used = _lib.MUTATE_FN(self._mutate_python), True
            a.mutation = _lib.MUTATION_NONE
        rd = self.reference_directions
        if isinstance(rd, PythonReferenceDirections) and hasattr(rd, "update"):
            h.update_directions, used = _lib.DIRECTIONS_FN(self._directions_python), True
        self._sel_state: dict[str, Any] = {}
        if self._python_selector:
            h.select, used = _lib.SELECT_FN(self._select_python), True
            h.prepare = _lib.PREPARE_FN(self._prepare_python)
        for strategy in (self.mating, self.crossover, self.mutation):
            if isinstance(strategy, (PythonMating, PythonCrossover, PythonMutation)) and hasattr(strategy, "initialize"):
                strategy.initialize(dict(info))
        if self._python_selector:
            # adapters.cpp:352-367
            sel_info = {key: info[key] for key in ("name", "population_size", "dimension", "objective_count",
                                                   "max_generations")}
            self.environment_selector.initialize(sel_info, self._directions.t(), {"device": self.cuda.device_id})
        self._hooks = h
        a.hooks = ctypes.pointer(h) if used else None

    @property
    def _python_selector(self) -> bool:
        return isinstance(self.environment_selector, PythonEnvironmentSelector)

    def _selector_start(self) -> None:
        """A Python selector's prepare() on the initial parents
        (algorithm.cu:196-206, adapters.cpp:369-375)."""
        self._sel_state = {}
        self._active.fill_(self.n)
        out = self.environment_selector.prepare(self._pop_view(self._current()), self._context(0, self.max_generations))
        self._update_state(out)
        self._check_mating_state()

    def _update_state(self, out: Any) -> None:
        # adapters.cpp:425-433: each key present replaces that part of the state
        if not isinstance(out, dict):
            return
        dev = self.cuda.device_id
        st = self._sel_state
        for key, dtype in (("rank", torch.int32), ("reference_index", torch.int32), ("score", torch.float32),
                           ("result_indices", torch.int32)):
            if key in out:
                st[key] = _checked(out[key], key, dev, dtype)
        if "active_count" in out:
            st["active_count"] = int(out["active_count"])
            self._active.fill_(st["active_count"])
        rank = st.get("rank")
        if rank is not None and rank.numel() == self.n:
            self._fronts[self._args.fronts_current].copy_(rank.reshape(self.n))

    def _check_mating_state(self) -> None:
        # mating_selection.cu: the tournament compares the selector's ranks
        rank = self._sel_state.get("rank")
        if isinstance(self.mating, TournamentMating) and (rank is None or rank.numel() != self.n):
            raise ValueError("TournamentMating requires a rank array for every parent")

    def _prepare_python(self, user, parents, n, d, m, generation, max_generations, stream) -> int:
        def body() -> int:
            # algorithm.cu:233-241: the parents changed, so the selector prepares again
            out = self.environment_selector.prepare(self._pop_dict(parents),
                                                    self._context(generation, max_generations))
            self._update_state(out)
            self._check_mating_state()
            return 0
        return self._guarded(body)

    def _select_python(self, user, parents, offspring, nxt, n, d, m, generation, max_generations, stream) -> int:
        def body() -> int:
            # adapters.cpp:377-404
            pd, od = self._pop_dict(parents), self._pop_dict(offspring)
            out = self.environment_selector.select(pd, od, self._context(generation, max_generations))
            if not isinstance(out, dict):
                raise ValueError("environment select must return a dict")
            dst = self._by_ptr[nxt.contents.x]
            if "indices" in out:
                idx = _checked(out["indices"], "selected indices", self.cuda.device_id, torch.int64)
                if idx.numel() != n:
                    raise ValueError("selected indices must contain population_size values")
                idx = idx.reshape(n)
                if bool(((idx < 0) | (idx >= 2 * n)).any()):
                    raise ValueError("selected indices must lie in [0, 2 * population_size)")
                dst.x.copy_(torch.cat([pd["variables"], od["variables"]]).index_select(0, idx))
                dst.f.copy_(torch.cat([pd["objectives"], od["objectives"]]).index_select(0, idx).t())
                dst.cv.copy_(torch.cat([pd["constraints"], od["constraints"]]).index_select(0, idx))
            else:
                x = _checked(out.get("variables"), "variables", self.cuda.device_id, torch.float32)
                if tuple(x.shape) != (n, d):
                    raise ValueError("variables must have shape (population_size, dimension)")
                dst.x.copy_(x)
                self._copy_evaluation(out, dst, n, m)
            self._update_state(out)
            if generation + 1 < max_generations:
                self._check_mating_state()
            return 0
        return self._guarded(body)

    def _copy_evaluation(self, out: Any, pop: K.Population, n: int, m: int) -> None:
        # adapters.cpp:115-145: objectives (n, m) and optional constraints (n)
        cons = None
        if isinstance(out, dict):
            obj, cons = out["objectives"], out.get("constraints")
        elif isinstance(out, tuple):
            obj = out[0]
            cons = out[1] if len(out) > 1 else None
        else:
            obj = out
        if not isinstance(obj, torch.Tensor) or not obj.is_cuda:
            raise ValueError("objectives must be a torch.Tensor on the algorithm CUDA device")
        if tuple(obj.shape) != (n, m):
            raise ValueError("objectives must have shape (population_size, objective_count)")
        pop.f.copy_(obj.to(torch.float32).t())
        if cons is not None:
            cons = torch.as_tensor(cons, device=self.device, dtype=torch.float32)
            if cons.numel() != n:
                raise ValueError("constraints must contain population_size values")
            pop.cv.copy_(cons.reshape(n))
        else:
            pop.cv.zero_()

    def _guarded(self, body) -> int:
        try:
            return body()
        except BaseException as exc:  # surfaced after the native call returns
            self._error = exc
            return -1

    def _mate_python(self, user, parents, n, d, m, rank, active_device, out, generation, max_generations,
                     stream) -> int:
        def body() -> int:
            # adapters.cpp:207-224
            state: dict[str, Any] = {}
            if self._python_selector:
                # adapters.cpp:412-417: the selector's mating state
                state.update({k: v for k, v in self._sel_state.items()
                              if k in ("rank", "reference_index", "score")})
            elif rank:
                state["rank"] = next(t for t in self._fronts if t.data_ptr() == rank)
            state["active_count"] = int(self._active.item()) if active_device else n
            idx = _checked(self.mating.select(self._pop_dict(parents), state,
                                              self._context(generation, max_generations)),
                           "parent_indices", self.cuda.device_id, torch.int32)
            if idx.numel() != n:
                raise ValueError("parent_indices has the wrong number of elements")
            idx = idx.reshape(n)
            if bool(((idx < 0) | (idx >= n)).any()):
                raise ValueError("parent_indices must lie in [0, population_size)")
            self._parent_idx.copy_(idx)
            return 0
        return self._guarded(body)

    def _crossover_python(self, user, parents, n, d, m, indices, offspring_x, generation, max_generations,
                          stream) -> int:
        def body() -> int:
            # adapters.cpp:238-251
            out = _checked(self.crossover.apply(self._pop_dict(parents), self._parent_idx,
                                                self._context(generation, max_generations)),
                           "offspring variables", self.cuda.device_id, torch.float32)
            if tuple(out.shape) != (n, d):
                raise ValueError("crossover output must have shape (population_size, dimension)")
            self._by_ptr[offspring_x].x.copy_(out)
            return 0
        return self._guarded(body)

    def _mutate_python(self, user, offspring, n, d, m, generation, max_generations, stream) -> int:
        def body() -> int:
            # adapters.cpp:268-281: None leaves the offspring as they are
            out = self.mutation.mutate(self._pop_dict(offspring), self._context(generation, max_generations))
            if out is not None:
                out = _checked(out, "mutated variables", self.cuda.device_id, torch.float32)
                if tuple(out.shape) != (n, d):
                    raise ValueError("mutation output must have shape (population_size, dimension)")
                self._by_ptr[offspring.contents.x].x.copy_(out)
            return 0
        return self._guarded(body)

    def _directions_python(self, user, nxt, n, d, m, active_device, directions, k, generation,
                           max_generations, stream) -> int:
        def body() -> int:
            # adapters.cpp:309-318, 325-335: the same count, (count, objective_count)
            active = int(self._active.item()) if active_device else n
            out = self.reference_directions.update(self._pop_dict(nxt), self._directions.t(), active,
                                                   self._context(generation, max_generations))
            if out is None:
                return 0
            out = _checked(out, "reference directions", self.cuda.device_id, torch.float32)
            if out.dim() != 2 or out.shape[1] != m:
                raise ValueError("reference directions must have shape (count, objective_count)")
            if out.shape[0] != k:
                raise ValueError("updated reference directions cannot change count")
            self._directions.copy_(out.t())
            return 1
        return self._guarded(body)

    def _evaluate_python(self, user, x, n, d, m, f, cv, generation, max_generations, stream) -> int:
        # adapters.cpp:115-173: evaluate on the run's stream, copy (n, m) -> (m, n)
        try:
            pop = self._by_ptr[x]
            out = self.problem.evaluate(pop.x, {"generation": generation,
                                                "max_generations": max_generations,
                                                "device": self.cuda.device_id})
            self._copy_evaluation(out, pop, n, m)
            return 0
        except BaseException as exc:  # surfaced after the native call returns
            self._error = exc
            return -1

    def _prepare_parent_python(self, user, parents, n, d, m, generation, max_generations, stream) -> int:
        # adapters.cpp:175-186: None / False -> unchanged; True -> changed in
        # place; an evaluation result replaces the parents' objectives and
        # constraints
        try:
            pop = self._by_ptr[parents.contents.x]
            out = self.problem.prepare_parent(self._pop_view(pop), {"generation": generation,
                                                                   "max_generations": max_generations,
                                                                   "device": self.cuda.device_id})
            if out is None:
                return 0
            if isinstance(out, bool):
                return 1 if out else 0
            self._copy_evaluation(out, pop, n, m)
            return 1
        except BaseException as exc:  # surfaced after the native call returns
            self._error = exc
            return -1

    def _call(self, fn, *extra) -> None:
        self._error = None
        status = fn(ctypes.byref(self._args), *extra)
        if self._error is not None:
            err, self._error = self._error, None
            raise err
        check(status)

    # --------------------------------------------------------- lifecycle ----
    def step(self) -> None:
        if not self._initialized:
            self.initialize()
        if self.finished:
            return
        self._call(lib.moea_run_step, 1)
        self._snapshot(self.generation)
        self._progress()

    def _snapshot(self, generation: int) -> None:
        # algorithm.cu:70-85: the current population, the current directions and
        # the selector's auxiliary indices (NSGA-III: the population's front
        # ranks; RVEA: the front-zero indices of the last result(), none before)
        rec = self._recorder
        if rec is None or not rec.should_save(generation, self.max_generations):
            return
        torch.cuda.current_stream(self.device).synchronize()
        pop = self._current()
        if self._python_selector:
            res = self._sel_state.get("result_indices")
            aux, active = (None if res is None else res.cpu().numpy()), self._sel_state.get("active_count", self.n)
        elif self.kind == "NSGA3":
            aux, active = self._fronts[self._args.fronts_current].cpu().numpy(), self.n
        else:
            aux, active = self._front0, int(self._active.item())
        rec.save(generation, pop.x.cpu().numpy(), pop.f.cpu().numpy(), pop.cv.cpu().numpy(),
                 self._directions.cpu().numpy(), aux, active)

    def _progress(self) -> None:
        g = self.generation
        if self.print_progress and (g == self.max_generations or
                                    (self.progress_interval > 0 and g % self.progress_interval == 0)):
            print(f"Generation {g} / {self.max_generations} completed")

    def run(self, *, copy: bool = True) -> Result:
        if not self._initialized:
            self.initialize()
        stream = torch.cuda.current_stream(self.device)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        while not self.finished:
            if self.print_progress and self.progress_interval > 0:
                todo = self.progress_interval - self.generation % self.progress_interval
            else:
                todo = self.max_generations - self.generation
            if self._recorder is not None:
                # stop at every generation the recorder saves (algorithm.cu:271-273)
                todo = min(todo, self._recorder.next_save(self.generation, self.max_generations) - self.generation)
            self._call(lib.moea_run_step, int(todo))
            self._snapshot(self.generation)
            self._progress()
        stop.record(stream)
        stop.synchronize()
        self._total_ms = float(start.elapsed_time(stop))
        return self.result(copy=copy)

    def _current(self) -> K.Population:
        return self._pops[self._args.current]

    def result(self, *, copy: bool = True) -> Result:
        if not self._initialized:
            raise RuntimeError("Algorithm has not been initialized")
        pop = self._current()
        torch.cuda.current_stream(self.device).synchronize()
        self._check_device()
        if self._python_selector:
            sel = self.environment_selector
            if hasattr(sel, "finalize"):         # adapters.cpp:405-411
                self._update_state(sel.finalize(self._pop_view(pop), self._context(self.generation,
                                                                                   self.max_generations)))
            aux = self._sel_state.get("result_indices")
            aux = torch.empty(0, dtype=torch.int32, device=self.device) if aux is None else aux
            active = self._sel_state.get("active_count", self.n)
        elif self.kind == "NSGA3":
            aux = self._fronts[self._args.fronts_current]
            active = self.n
        else:
            active = int(self._active.item())
            fr, _ = K.nds_rank(pop.f[:, :active].contiguous(), pop.cv[:active].contiguous())
            aux = torch.nonzero(fr == 0).flatten().to(torch.int32)
            self._front0 = aux.cpu().numpy()
        self._snapshot(self.generation)          # algorithm.cu:336-339, after finalize
        x, f, cv = pop.x, pop.f.t(), pop.cv
        if copy:
            x, f, cv, aux = x.clone(), f.contiguous(), cv.clone(), aux.clone()
        return Result(x, f, cv, aux, active, self._total_ms)

    def reset(self) -> None:
        if not self._initialized:
            return
        torch.cuda.current_stream(self.device).synchronize()
        self._stats.zero_()                      # a new run starts with a clear device-error word
        self._resets += 1
        self._directions.copy_(self._directions0)
        rd = self.reference_directions
        if isinstance(rd, PythonReferenceDirections) and hasattr(rd, "reset"):
            rd.reset()                           # adapters.cpp:319-324
        if self._native is not None:
            self._native.reset()                 # algorithm.cu:413, native_problem.cpp:142
        elif self._args.problem.kind == _lib.PROBLEM_CALLBACK and hasattr(self.problem, "reset"):
            self.problem.reset()                 # algorithm.cu:413, adapters.cpp:187-190
        self._args.seed = self._run_seed()
        self._args.current = 0
        self._args.fronts_current = 0
        self._args.initial_x_given = 0
        if self._python_selector and hasattr(self.environment_selector, "reset"):
            self.environment_selector.reset()    # adapters.cpp:412-420
        self._call(lib.moea_run_init)
        if self._python_selector:
            self._selector_start()
        self._total_ms = 0.0
        self._front0 = None
        self._snapshot(0)                        # algorithm.cu:430

    def synchronize(self) -> None:
        if self._initialized:
            torch.cuda.current_stream(self.device).synchronize()
            self._check_device()

    def _check_device(self) -> None:
        """Raises if a front-peeling launch of this run gave up waiting at its
        grid barrier (the sticky device-error stats word; read after a
        synchronization the caller already made)."""
        if self.kind == "NSGA3" and int(self._stats[_lib.STAT_DEVICE_ERROR].item()) != 0:
            raise MoeaError("a front-peeling launch gave up waiting at its grid barrier; the selections of "
                            "the affected generations kept their parents")

    @property
    def population(self) -> Population:
        if not self._initialized:
            raise RuntimeError("Algorithm has not been initialized")
        pop = self._current()
        return Population(pop.x, pop.f.t(), pop.cv)

    @property
    def initialized(self) -> bool:
        return self._initialized

    @property
    def finished(self) -> bool:
        return self._initialized and self._args.generation >= self.max_generations

    @property
    def generation(self) -> int:
        return int(self._args.generation) if self._initialized else 0

    @property
    def directions(self) -> torch.Tensor:
        return self._directions


class NSGA3(Algorithm):
    def __init__(self, **kwargs: Any) -> None:
        super().__init__("NSGA3", **kwargs)


class RVEA(Algorithm):
    def __init__(self, **kwargs: Any) -> None:
        super().__init__("RVEA", **kwargs)


class AlgorithmBuilder:
    def __init__(self, algorithm: str) -> None:
        self.algorithm = algorithm
        self.options: dict[str, Any] = {}

    def set(self, **options: Any) -> "AlgorithmBuilder":
        self.options.update(options)
        return self

    def population_size(self, value: int) -> "AlgorithmBuilder":
        return self.set(population_size=value)

    def max_generations(self, value: int) -> "AlgorithmBuilder":
        return self.set(max_generations=value)

    def problem(self, value: Any) -> "AlgorithmBuilder":
        return self.set(problem=value)

    def mating(self, value: Any) -> "AlgorithmBuilder":
        return self.set(mating=value)

    def crossover(self, value: Any) -> "AlgorithmBuilder":
        return self.set(crossover=value)

    def mutation(self, value: Any) -> "AlgorithmBuilder":
        return self.set(mutation=value)

    def reference_directions(self, value: Any) -> "AlgorithmBuilder":
        return self.set(reference_directions=value)

    def environment_selector(self, value: Any) -> "AlgorithmBuilder":
        return self.set(environment_selector=value)

    def build(self) -> Algorithm:
        return Algorithm(self.algorithm, **self.options)


def _device_index(device: int | str) -> int:
    if isinstance(device, int):
        return device
    if device == "cuda":
        return 0
    if device.startswith("cuda:"):
        return int(device.split(":", 1)[1])
    raise ValueError("cuda_moea only supports GPU devices")
