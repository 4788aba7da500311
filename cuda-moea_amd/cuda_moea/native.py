"""Native (C++/HIP) implementations of IProblemEvaluator loaded as plugins.

The Python face of the reference's native problems (python/cuda_moea/
native.py:13-83): ``NativeProblem(library=..., name=..., dimension=...,
objectives=..., parameters=...)`` validates a plugin library's registration
and parameters; an algorithm built on it creates the problem instance through
libmoea_hip's loader (moea_plugin_*), and every evaluation is a C call from
the run loop to the plugin, with no Python on that path.  A plugin is built
with hipcc against include/cuda_moea/native/plugin.hpp (the Makefile's
lib/plugins/libbi_sphere.so is an example); building from a source folder
through a compile cache is not part of this package.
"""
from __future__ import annotations

import copy
import ctypes
import math
import struct
import time
import weakref
from pathlib import Path

from . import _lib
from ._lib import check, lib

_VALIDATORS = {
    "int": lambda v: type(v) is int and -(2 ** 63) <= v < 2 ** 63,
    "float": lambda v: type(v) in (int, float) and math.isfinite(v),
    "bool": lambda v: type(v) is bool,
    "str": lambda v: isinstance(v, str),
    "float_list": lambda v: isinstance(v, (list, tuple)) and all(
        type(x) in (int, float) and math.isfinite(x) for x in v),
}


class _Plugin:
    """An open plugin library (moea_plugin_open); closed when collected."""

    def __init__(self, path: str):
        h = ctypes.c_void_p()
        check(lib.moea_plugin_open(str(path).encode(), ctypes.byref(h)))
        self.handle = h.value
        self._close = weakref.finalize(self, lib.moea_plugin_close, h.value)

    def schema(self) -> dict:
        params = {}
        for i in range(lib.moea_plugin_param_count(self.handle)):
            v, desc = _lib.ParamValue(), ctypes.c_char_p()
            check(lib.moea_plugin_param(self.handle, i, ctypes.byref(v), ctypes.byref(desc)))
            kind = _lib.PARAM_TYPES[v.type]
            default = {"int": lambda: int(v.int_value), "float": lambda: float(v.float_value),
                       "bool": lambda: bool(v.int_value), "str": lambda: v.str_value.decode(),
                       "float_list": lambda: [float(v.list[j]) for j in range(v.list_length)]}[kind]()
            params[v.name.decode()] = {"type": kind, "default": default,
                                       "description": (desc.value or b"").decode()}
        return {"name": lib.moea_plugin_name(self.handle).decode(),
                "build_id": lib.moea_plugin_build_id(self.handle).decode(), "parameters": params}


def native_problem_schema(path) -> dict:
    """Registered name, build id and parameter schema of a plugin library
    (native_problem.cpp:155-170)."""
    return _Plugin(str(path)).schema()


def _param_values(parameters: dict) -> tuple:
    """ctypes moea_param_value entries for `parameters`, typed by their Python
    type; the loader checks them against the schema again."""
    keep, out = [], (_lib.ParamValue * max(len(parameters), 1))()
    for i, (key, value) in enumerate(parameters.items()):
        pv = out[i]
        pv.name = str(key).encode()
        keep.append(pv.name)
        if isinstance(value, bool):
            pv.type, pv.int_value = _lib.PARAM_BOOL, int(value)
        elif isinstance(value, int):
            if not -(2 ** 63) <= value < 2 ** 63:
                raise ValueError(f"Invalid type/value for native parameter '{key}'")
            pv.type, pv.int_value = _lib.PARAM_INT, value
        elif isinstance(value, float):
            pv.type, pv.float_value = _lib.PARAM_FLOAT, value
        elif isinstance(value, str):
            pv.type, pv.str_value = _lib.PARAM_STR, value.encode()
            keep.append(pv.str_value)
        elif isinstance(value, (list, tuple)):
            if not all(type(x) in (int, float) for x in value):
                raise ValueError(f"Invalid type/value for native parameter '{key}'")
            arr = (ctypes.c_double * max(len(value), 1))(*[float(x) for x in value])
            keep.append(arr)
            pv.type, pv.list_length = _lib.PARAM_FLOAT_LIST, len(value)
            pv.list = ctypes.cast(arr, ctypes.POINTER(ctypes.c_double))
        else:
            raise ValueError(f"Invalid type/value for native parameter '{key}'")
    return out, keep


class _Instance:
    """One created problem (moea_plugin_create) and its C callbacks."""

    def __init__(self, spec: dict):
        self._plugin = _Plugin(spec["library"])
        if self._plugin.schema()["name"] != spec["name"]:
            raise ValueError("Native problem name does not match registration")
        d, m = int(spec["dimension"]), int(spec["objectives"])
        lb = (ctypes.c_float * d)(*spec["lower_bounds"])
        ub = (ctypes.c_float * d)(*spec["upper_bounds"])
        params, _keep = _param_values(spec["parameters"])
        h = ctypes.c_void_p()
        check(lib.moea_plugin_create(self._plugin.handle, d, m, lb, ub, params, len(spec["parameters"]),
                                     ctypes.byref(h)))
        self.handle = h.value
        self.constraints = lib.moea_plugin_problem_constraints(self.handle)
        self._destroy = weakref.finalize(self, lib.moea_plugin_problem_destroy, h.value)

    def initialize(self, stream: int, seed: int, bounds_ptr: int) -> None:
        check(lib.moea_plugin_problem_initialize(self.handle, stream, seed, bounds_ptr))

    def reset(self) -> None:
        check(lib.moea_plugin_reset(self.handle))

    def destroy(self) -> None:
        self._destroy()


# the loader's C entry points as the run loop's callback pointers
EVALUATE = _lib.EVALUATE_FN(ctypes.cast(lib.moea_plugin_evaluate, ctypes.c_void_p).value)
PREPARE_PARENT = _lib.PREPARE_PARENT_FN(ctypes.cast(lib.moea_plugin_prepare_parent, ctypes.c_void_p).value)


class NativeProblem:
    """A problem implemented by a plugin library (native.py:23-83 of the
    reference): the same keyword arguments, bound and parameter checks and
    ``schema`` / ``_spec()``.  ``source_dir`` builds are not supported: pass
    ``library=`` the plugin's path."""

    def __init__(self, source_dir=None, *, name: str, dimension: int, objectives: int, lower_bounds=0.0,
                 upper_bounds=1.0, parameters=None, library=None, build="auto", **build_options):
        if (source_dir is None) == (library is None):
            raise ValueError("Provide exactly one of source_dir or library")
        if build not in {"auto", "never"}:
            raise ValueError("build must be 'auto' or 'never'")
        for key, value in (("dimension", dimension), ("objectives", objectives)):
            if isinstance(value, bool) or not isinstance(value, int) or value <= 0:
                raise ValueError(f"{key} must be a positive integer")
        if not isinstance(name, str) or not name:
            raise ValueError("name must be a nonempty string")
        lower = self._bounds(lower_bounds, dimension)
        upper = self._bounds(upper_bounds, dimension)
        if any(a >= b for a, b in zip(lower, upper)):
            raise ValueError("Each lower bound must be less than its upper bound")
        if library is not None and build_options:
            raise ValueError("Build options are only applicable to source_dir")
        if library is None:
            raise NotImplementedError("building a plugin from source_dir is not supported; build it with hipcc "
                                      "against include/cuda_moea/native/plugin.hpp and pass library=")
        self.library = Path(library).expanduser().resolve()
        if not self.library.is_file():
            raise FileNotFoundError(self.library)
        self.schema = native_problem_schema(self.library)
        if self.schema["name"] != name:
            raise ValueError(f"Library registers {self.schema['name']!r}, not {name!r}")
        self._config = {"type": "NativeProblem", "library": str(self.library), "name": name,
                        "dimension": dimension, "objectives": objectives, "lower_bounds": lower,
                        "upper_bounds": upper, "parameters": copy.deepcopy(dict(parameters or {}))}
        self._validate_parameters()

    @staticmethod
    def _bounds(value, dimension):
        values = [float(value)] * dimension if isinstance(value, (int, float)) else list(map(float, value))
        if len(values) != dimension or not all(math.isfinite(x) and abs(x) <= 3.4028234663852886e38
                                               for x in values):
            raise ValueError("Bounds must be finite float32 values with length dimension")
        return [struct.unpack("f", struct.pack("f", x))[0] for x in values]

    def _validate_parameters(self):
        for key, value in self._config["parameters"].items():
            if key not in self.schema["parameters"]:
                raise ValueError(f"Unknown native problem parameter: {key}")
            kind = self.schema["parameters"][key]["type"]
            if not _VALIDATORS[kind](value):
                raise ValueError(f"Native parameter {key!r} must be {kind}")

    def _spec(self) -> dict:
        return copy.deepcopy(self._config)

    def _instantiate(self) -> _Instance:
        return _Instance(self._spec())

    # the attributes the algorithms read from a problem
    @property
    def name(self) -> str:
        return self._config["name"]

    @property
    def dimension(self) -> int:
        return self._config["dimension"]

    @property
    def objectives(self) -> int:
        return self._config["objectives"]

    @property
    def lower_bounds(self) -> list:
        return list(self._config["lower_bounds"])

    @property
    def upper_bounds(self) -> list:
        return list(self._config["upper_bounds"])


def benchmark(problem: NativeProblem, variables, *, repeats: int = 100, warmup: int = 10) -> dict:
    """Time whole-batch evaluations of `variables` (N, D) on the device
    (native.py:13-20 of the reference): HIP events give the stream timeline
    per evaluation, wall time adds the host's launches and the final wait."""
    import torch
    if repeats <= 0 or warmup < 0:
        raise ValueError("repeats must be positive and warmup nonnegative")
    x = torch.as_tensor(variables)
    if x.dim() != 2 or x.shape[1] != problem.dimension or x.dtype != torch.float32 or not x.is_cuda:
        raise ValueError("variables must be a float32 device tensor of shape (N, dimension)")
    x = x.contiguous()
    n, d, m = x.shape[0], problem.dimension, problem.objectives
    inst = problem._instantiate()
    dev = x.device
    stream = torch.cuda.current_stream(dev)
    bounds = torch.tensor([v for pair in zip(problem.lower_bounds, problem.upper_bounds) for v in pair],
                          dtype=torch.float32, device=dev)
    f = torch.empty(m, n, dtype=torch.float32, device=dev)
    cv = torch.empty(n, dtype=torch.float32, device=dev)
    inst.initialize(stream.cuda_stream, 0, bounds.data_ptr())

    def once():
        check(lib.moea_plugin_evaluate(inst.handle, x.data_ptr(), n, d, m, f.data_ptr(), cv.data_ptr(), 0, 1,
                                       stream.cuda_stream))

    for _ in range(warmup):
        once()
    stream.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(repeats):
        once()
    stop.record(stream)
    stop.synchronize()
    wall = (time.perf_counter() - t0) * 1e3
    inst.destroy()
    return {"wall_ms_per_eval": wall / repeats, "cuda_ms_per_eval": start.elapsed_time(stop) / repeats,
            "repeats": repeats, "warmup": warmup}
