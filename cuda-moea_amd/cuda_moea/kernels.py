"""Stage-level entry points of libmoea_hip.so on torch device tensors.

Each function validates shapes on the host, allocates its outputs and the
workspace with torch (torch is the allocator and the stream provider only) and
calls one C-ABI function on the current stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from ._lib import DEVICE_FAILED, MoeaError, check, lib

_F32, _I32 = torch.float32, torch.int32


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _dev(t: torch.Tensor, dtype, name: str, shape=None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    return t


def workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


@dataclass
class Population:
    """Device population in the reference's layouts: x (n, d), f (m, n), cv (n,)."""
    x: torch.Tensor
    f: torch.Tensor
    cv: torch.Tensor

    def c(self) -> _lib.PopulationC:
        return _lib.PopulationC(self.x.data_ptr(), self.f.data_ptr(), self.cv.data_ptr())

    @staticmethod
    def empty(n: int, d: int, m: int, device) -> "Population":
        return Population(torch.empty(n, d, dtype=_F32, device=device),
                          torch.empty(m, n, dtype=_F32, device=device),
                          torch.empty(n, dtype=_F32, device=device))


_DRAW_TYPES = {"mate_a": _I32, "mate_b": _I32, "sbx_u": _F32, "sbx_p": _F32, "sbx_sign": _I32,
               "sbx_copy": _F32, "pm_gate": _F32, "pm_delta": _F32, "niche_a": torch.uint32,
               "niche_b": torch.uint32}


def make_draws(draws: Optional[dict], device) -> tuple[Optional[_lib.Draws], list]:
    """Device copies of injected draws; missing keys are generated in-kernel."""
    if draws is None:
        return None, []
    keep = []
    st = _lib.Draws()
    for key, dtype in _DRAW_TYPES.items():
        if key in draws and draws[key] is not None:
            v = draws[key]
            if isinstance(v, np.ndarray):
                v = torch.from_numpy(np.ascontiguousarray(v))
            if dtype == torch.uint32:
                t = v.to(device=device).contiguous().view(torch.int32) if v.dtype == torch.uint32 \
                    else torch.as_tensor(v, dtype=torch.int64).to(torch.int32).to(device).contiguous()
            else:
                t = v.to(device=device, dtype=dtype).contiguous()
            keep.append(t)
            setattr(st, key, t.data_ptr())
    return st, keep


# ------------------------------------------------------------------ sorting --
def nds_rank(f: torch.Tensor, cv: torch.Tensor, stop_at: Optional[int] = None):
    """Constrained NDS of the columns of f (m, n).  Returns (fronts (n,) int32
    device tensor, max_front int)."""
    _dev(f, _F32, "f")
    m, n = f.shape
    _dev(cv, _F32, "cv", (n,))
    fronts = torch.empty(n, dtype=_I32, device=f.device)
    mf = torch.empty(1, dtype=_I32, device=f.device)
    ws = workspace(lib.moea_nds_workspace_bytes(n), f.device)
    check(lib.moea_nds_rank(f.data_ptr(), n, cv.data_ptr(), n, m, n if stop_at is None else stop_at,
                            fronts.data_ptr(), mf.data_ptr(), ws.data_ptr(), ws.numel(),
                            _stream(f.device)))
    max_front = int(mf.item())
    if max_front == DEVICE_FAILED:
        raise MoeaError("nds_rank: the front-peeling launch gave up waiting at its grid barrier; "
                        "the fronts are invalid")
    return fronts, max_front


# -------------------------------------------------------------- directions --
def das_dennis(requested_count: int, m: int, partitions: int = 0, kind: Optional[int] = None,
               raw: bool = False):
    """Das-Dennis unit directions (m, K) as a numpy array (host-side, as in the
    reference).  kind: 0 by count, 1 by partitions, 2 RVEA by count."""
    if kind is None:
        kind = 1 if partitions > 0 else 0
    k = lib.moea_das_dennis(requested_count, m, partitions, kind, None, None)
    if k <= 0:
        check(k if k < 0 else -1)
    out = np.zeros((m, k), np.float32)
    rw = np.zeros((m, k), np.float32)
    check(min(0, lib.moea_das_dennis(requested_count, m, partitions, kind,
                                     out.ctypes.data_as(ctypes.c_void_p),
                                     rw.ctypes.data_as(ctypes.c_void_p))))
    return (out, rw) if raw else out


# -------------------------------------------------------------- selection --
def ideal_point(f: torch.Tensor) -> torch.Tensor:
    _dev(f, _F32, "f")
    m, n = f.shape
    z = torch.empty(m, dtype=_F32, device=f.device)
    check(lib.moea_ideal_point(f.data_ptr(), n, n, m, z.data_ptr(), _stream(f.device)))
    return z


def associate(fn: torch.Tensor, a: torch.Tensor, directions: torch.Tensor, method: int = 0):
    """Association of normalized objectives fn (m, n) scaled by intercepts a (m,)
    with unit directions (m, k): (dist (n,), ref (n,)).  method 0 automatic,
    1 every direction, 2 cone pruning (identical results)."""
    _dev(fn, _F32, "fn")
    m, n = fn.shape
    _dev(a, _F32, "a", (m,))
    _dev(directions, _F32, "directions")
    k = directions.shape[1]
    dist = torch.empty(n, dtype=_F32, device=fn.device)
    ref = torch.empty(n, dtype=_I32, device=fn.device)
    ws = workspace(lib.moea_associate_workspace_bytes(n, m, k), fn.device)
    check(lib.moea_associate(fn.data_ptr(), n, n, m, a.data_ptr(), directions.data_ptr(), k, dist.data_ptr(),
                             ref.data_ptr(), int(method), ws.data_ptr(), ws.numel(), _stream(fn.device)))
    return dist, ref


def cv_quantization(bins: int = 0, clip_upper: float = 0.0, log_alpha: float = 0.0,
                    feasibility_epsilon: float = 0.0) -> _lib.CvQuantization:
    return _lib.CvQuantization(int(bins), float(clip_upper), float(log_alpha), float(feasibility_epsilon))


def quantize_cv(cv: torch.Tensor, **params) -> torch.Tensor:
    """The sort's bins of the violations cv (n,): 0 feasible, else 1..bins."""
    _dev(cv, _F32, "cv")
    out = torch.empty_like(cv)
    q = cv_quantization(**params)
    check(lib.moea_quantize_cv(cv.data_ptr(), out.data_ptr(), cv.numel(), ctypes.byref(q), _stream(cv.device)))
    return out


def nsga3_select(parents: Population, offspring: Population, directions: torch.Tensor,
                 ideal: torch.Tensor, draws: Optional[dict] = None, seed: int = 0,
                 generation: int = 0, cv_quant: Optional[dict] = None, normalization: Optional[dict] = None,
                 niching: tuple = (0, 0)):
    """One NSGA-III selection.  ``ideal`` is updated in place.  Returns
    (next Population, next_fronts, stats int32[8] = max_front, n_prior, n_last,
    fallback, n_rem, ...).  With a ``normalization`` dict, its "extremes"
    (int32 (m,)) and "intercepts" (float64 (m,)) device tensors receive the
    selection's extreme points and intercepts when it normalizes.  ``niching``
    = (mode, lds_cap) of moea_niching ((0, 0): automatic)."""
    m, n = parents.f.shape
    d = parents.x.shape[1]
    k = directions.shape[1]
    dev = parents.x.device
    for name, p in (("parents", parents), ("offspring", offspring)):
        _dev(p.x, _F32, f"{name}.x", (n, d)); _dev(p.f, _F32, f"{name}.f", (m, n))
        _dev(p.cv, _F32, f"{name}.cv", (n,))
    _dev(directions, _F32, "directions", (m, k)); _dev(ideal, _F32, "ideal", (m,))
    nxt = Population.empty(n, d, m, dev)
    nfr = torch.empty(n, dtype=_I32, device=dev)
    stats = torch.zeros(8, dtype=_I32, device=dev)
    ws = workspace(lib.moea_nsga3_workspace_bytes(n, m, k), dev)
    dr, keep = make_draws(draws, dev)
    a = _lib.NSGA3SelectArgs()
    a.n, a.m, a.d, a.k = n, m, d, k
    a.parents, a.offspring, a.next = parents.c(), offspring.c(), nxt.c()
    a.next_fronts, a.directions, a.ideal = nfr.data_ptr(), directions.data_ptr(), ideal.data_ptr()
    a.draws = ctypes.pointer(dr) if dr is not None else None
    a.seed, a.generation = seed, generation
    a.niching.mode, a.niching.lds_cap = int(niching[0]), int(niching[1])
    a.stats, a.workspace, a.workspace_bytes = stats.data_ptr(), ws.data_ptr(), ws.numel()
    a.stream = _stream(dev)
    if cv_quant:
        a.cv_quant = cv_quantization(**cv_quant)
    if normalization is not None:
        ext = normalization.setdefault("extremes", torch.full((m,), -7, dtype=_I32, device=dev))
        icp = normalization.setdefault("intercepts", torch.full((m,), -7.0, dtype=torch.float64, device=dev))
        _dev(ext, _I32, "normalization extremes", (m,))
        if icp.dtype != torch.float64 or icp.device != dev or tuple(icp.shape) != (m,):
            raise ValueError("normalization intercepts must be a float64 (m,) tensor on the device")
        a.out_extremes, a.out_intercepts = ext.data_ptr(), icp.data_ptr()
    check(lib.moea_nsga3_select(ctypes.byref(a)))
    return nxt, nfr, stats


# ----------------------------------------------------------- reproduction --
def reproduce(parents_x: torch.Tensor, bounds: torch.Tensor, *, rank=None, cv=None,
              mating: int = _lib.MATING_TOURNAMENT, active: Optional[int] = None,
              active_device: Optional[torch.Tensor] = None,
              sbx=(30.0, 1.0, 0.0), mutation: int = _lib.MUTATION_POLYNOMIAL, pm=(20.0, 1.0),
              draws: Optional[dict] = None, seed: int = 0, generation: int = 0):
    """Mating -> SBX -> PM.  sbx = (eta, probability, copy probability),
    pm = (eta, probability).  Candidates come from the first ``active`` parents,
    or from the first active_device[0] (an int32 device tensor, read by the
    kernel) when given.  Returns (offspring_x, parent_indices)."""
    n, d = parents_x.shape
    dev = parents_x.device
    _dev(parents_x, _F32, "parents_x"); _dev(bounds, _F32, "bounds", (2 * d,))
    off = torch.empty(n, d, dtype=_F32, device=dev)
    idx = torch.empty(n, dtype=_I32, device=dev)
    dr, keep = make_draws(draws, dev)
    a = _lib.ReproduceArgs()
    a.n, a.d, a.active = n, d, n if active is None else active
    if active_device is not None:
        _dev(active_device, _I32, "active_device", (1,))
        a.active_device = active_device.data_ptr()
    a.mating = mating
    a.parents_x, a.parents_rank, a.parents_cv = parents_x.data_ptr(), _ptr(rank), _ptr(cv)
    a.bounds = bounds.data_ptr()
    a.sbx_eta, a.sbx_prob, a.sbx_copy_prob = sbx
    a.mutation = mutation
    a.pm_eta, a.pm_prob = pm
    a.draws = ctypes.pointer(dr) if dr is not None else None
    a.seed, a.generation = seed, generation
    a.offspring_x, a.parent_indices, a.stream = off.data_ptr(), idx.data_ptr(), _stream(dev)
    check(lib.moea_reproduce(ctypes.byref(a)))
    return off, idx


def init_population(n: int, d: int, bounds: torch.Tensor, seed: int) -> torch.Tensor:
    x = torch.empty(n, d, dtype=_F32, device=bounds.device)
    check(lib.moea_init_population(x.data_ptr(), n, d, bounds.data_ptr(), seed,
                                   _stream(bounds.device)))
    return x


# -------------------------------------------------------------- objectives --
def eval_quadratic(x: torch.Tensor, centers: torch.Tensor):
    _dev(x, _F32, "x")
    n, d = x.shape
    _dev(centers, _F32, "centers")
    m = centers.shape[0]
    f = torch.empty(m, n, dtype=_F32, device=x.device)
    cv = torch.empty(n, dtype=_F32, device=x.device)
    check(lib.moea_eval_quadratic(x.data_ptr(), n, d, centers.data_ptr(), m, f.data_ptr(), n,
                                  cv.data_ptr(), _stream(x.device)))
    return f, cv


def eval_dtlz(kind: int, x: torch.Tensor, m: int):
    """DTLZ1-DTLZ7 (kind 1-7) or ConvexDTLZ2 (kind 8) objectives (m, n) and zero cv of x (n, d)."""
    _dev(x, _F32, "x")
    n, d = x.shape
    f = torch.empty(m, n, dtype=_F32, device=x.device)
    cv = torch.empty(n, dtype=_F32, device=x.device)
    check(lib.moea_eval_dtlz(int(kind), x.data_ptr(), n, d, m, f.data_ptr(), n, cv.data_ptr(),
                             _stream(x.device)))
    return f, cv


def eval_mlp(x: torch.Tensor, w1: torch.Tensor, w2: torch.Tensor):
    _dev(x, _F32, "x")
    n, d = x.shape
    hidden, m = w1.shape[0], w2.shape[0]
    f = torch.empty(m, n, dtype=_F32, device=x.device)
    cv = torch.empty(n, dtype=_F32, device=x.device)
    h = torch.empty(n, hidden, dtype=_F32, device=x.device)
    check(lib.moea_eval_mlp(x.data_ptr(), n, d, w1.data_ptr(), hidden, w2.data_ptr(), m,
                            f.data_ptr(), n, cv.data_ptr(), h.data_ptr(), _stream(x.device)))
    return f, cv


# -------------------------------------------------------------------- RVEA --
def rvea_gamma(directions: torch.Tensor) -> torch.Tensor:
    _dev(directions, _F32, "directions")
    m, k = directions.shape
    g = torch.empty(k, dtype=_F32, device=directions.device)
    check(lib.moea_rvea_gamma(directions.data_ptr(), m, k, g.data_ptr(),
                              _stream(directions.device)))
    return g


def rvea_penalty(m: int, t: int, max_generations: int, alpha: float) -> float:
    return float(lib.moea_rvea_penalty(m, t, max_generations, alpha))


def rvea_select(parents: Population, offspring: Population, directions: torch.Tensor,
                gamma: torch.Tensor, active: int, penalty: float):
    """Returns (next Population, new active count)."""
    m, n = parents.f.shape
    d = parents.x.shape[1]
    k = directions.shape[1]
    dev = parents.x.device
    nxt = Population.empty(n, d, m, dev)
    out_active = torch.empty(1, dtype=_I32, device=dev)
    ws = workspace(lib.moea_rvea_workspace_bytes(n, m, k), dev)
    a = _lib.RVEASelectArgs()
    a.n, a.m, a.d, a.k, a.active = n, m, d, k, active
    a.active_device = None
    a.parents, a.offspring, a.next = parents.c(), offspring.c(), nxt.c()
    a.directions, a.gamma, a.penalty = directions.data_ptr(), gamma.data_ptr(), penalty
    a.out_active = out_active.data_ptr()
    a.workspace, a.workspace_bytes, a.stream = ws.data_ptr(), ws.numel(), _stream(dev)
    check(lib.moea_rvea_select(ctypes.byref(a)))
    return nxt, int(out_active.item())


def rvea_adapt(v0: torch.Tensor, f: torch.Tensor, n_active: int) -> torch.Tensor:
    m, k = v0.shape
    dev = v0.device
    out = torch.empty(m, k, dtype=_F32, device=dev)
    na = torch.tensor([n_active], dtype=_I32, device=dev)
    scratch = torch.empty(m, dtype=_F32, device=dev)
    check(lib.moea_rvea_adapt(v0.data_ptr(), f.data_ptr(), f.shape[1], na.data_ptr(), m, k,
                              out.data_ptr(), scratch.data_ptr(), _stream(dev)))
    return out
