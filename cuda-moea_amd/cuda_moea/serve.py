"""Continuous batching of NSGA-III runs (libmoea_hip's ``moea_serve_*``).

A ``Server`` holds up to ``max_instances`` runs at once inside one device
workspace of ``budget_bytes`` (the HBM budget).  Runs arrive with ``admit``
while others are in flight; every ``step`` initializes the runs admitted since
the last one and advances every running run by one generation with a single
batched launch sequence; runs that reach their ``max_generations`` are
reported as finished, read with ``result`` and freed with ``release``.  Each
run computes exactly what the same ``NSGA3`` computes alone, whatever else
shares the steps.  The reference runs one population per EvolutionaryAlgorithm
(NS/algorithm/algorithm.cu:104-329) and has no serving loop.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import check, lib
from . import kernels as K
from .algorithm import NSGA3, Result
from .batch import check_run, fill_problem, view_result


class Server:
    def __init__(self, budget_bytes: int, *, max_instances: int = 256, max_rows: int | None = None,
                 device: int | torch.device = 0, niching: tuple[int, int] = (0, 0), graph: bool = False) -> None:
        """``graph``: steps between changes of the set of runs replay one
        captured generation (moea_serve_set_graph)."""
        if not torch.cuda.is_available():
            raise RuntimeError("cuda_moea requires a ROCm GPU")
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        cfg = _lib.ServeConfig()
        cfg.max_instances = int(max_instances)
        cfg.max_rows = int(max_rows if max_rows is not None else 2 * 16384 * max_instances)
        cfg.niching.mode, cfg.niching.lds_cap = int(niching[0]), int(niching[1])
        reserved = lib.moea_serve_reserved_bytes(ctypes.byref(cfg))
        if reserved == 0:
            raise ValueError("invalid server configuration")
        if budget_bytes <= reserved:
            raise ValueError(f"budget of {budget_bytes} bytes is below the {reserved} bytes the server reserves")
        self._cfg = cfg
        self._ws = K.workspace(int(budget_bytes), self.device)
        self._handle = ctypes.c_void_p()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(lib.moea_serve_create(ctypes.byref(cfg), self._ws.data_ptr(), self._ws.numel(), stream,
                                    ctypes.byref(self._handle)))
        check(lib.moea_serve_set_graph(self._handle, 1 if graph else 0))
        self._runs: dict[int, NSGA3] = {}
        self._keep: dict[int, list] = {}
        self._finished_buf = (ctypes.c_int32 * cfg.max_instances)()

    @staticmethod
    def instance_bytes(run: NSGA3) -> int:
        """Arena bytes ``run`` takes (its problem arrays are the caller's)."""
        check_run(run)
        e = _lib.BatchProblem()
        keep: list = []
        fill_problem(e, run, torch.device("cpu"), keep)
        return int(lib.moea_serve_instance_bytes(ctypes.byref(e)))

    def admit(self, run: NSGA3) -> int | None:
        """Place ``run`` in a free slot; None when no slot, row capacity or
        budget is free (retry after a release)."""
        check_run(run)
        if run.cuda.device_id != self.device.index:
            raise ValueError("run and server must use the same device")
        e = _lib.BatchProblem()
        keep: list = []
        fill_problem(e, run, self.device, keep)
        slot = ctypes.c_int(-1)
        st = lib.moea_serve_admit(self._handle, ctypes.byref(e), ctypes.byref(slot))
        if st == -3:
            return None
        check(st)
        self._runs[slot.value] = run
        self._keep[slot.value] = keep
        return slot.value

    def step(self) -> list[int]:
        """Initialize new runs and advance every running run by one
        generation; returns the slots of the runs that finished."""
        n = ctypes.c_int(0)
        check(lib.moea_serve_step(self._handle, self._finished_buf, self._cfg.max_instances, ctypes.byref(n)))
        return [int(self._finished_buf[i]) for i in range(n.value)]

    def generation(self, slot: int) -> int:
        g = lib.moea_serve_generation(self._handle, int(slot))
        check(min(g, 0))
        return g

    def result(self, slot: int, *, copy: bool = True) -> Result:
        pop = _lib.PopulationC()
        fronts = ctypes.c_void_p()
        check(lib.moea_serve_view(self._handle, int(slot), ctypes.byref(pop), ctypes.byref(fronts)))
        # waits for the queued generations and raises if one of them failed on the device
        check(lib.moea_serve_check(self._handle))
        return view_result(self._ws, pop, fronts.value, self._runs[slot], 0.0, copy)

    def release(self, slot: int) -> None:
        check(lib.moea_serve_release(self._handle, int(slot)))
        self._runs.pop(slot, None)
        self._keep.pop(slot, None)

    @property
    def live(self) -> list[int]:
        return sorted(self._runs)

    def __del__(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            lib.moea_serve_destroy(h)
            self._handle = ctypes.c_void_p()
