"""Run snapshots on disk, in the reference's format
(python/csrc/native/src/algorithm/data_recorder.cu:129-290).

Layout under the root directory:

* ``lower_bounds.bin``, ``upper_bounds.bin``: float32 (D)
* ``initial_reference_directions.bin``: float32 (M, K) objective-major
* ``metadata.json``: run description (format_version 1)
* ``generation_%06d/``: ``variables.bin`` float32 (N, D) row-major,
  ``objectives.bin`` float32 (M, N) objective-major, ``constraints.bin``
  float32 (N), ``reference_directions.bin`` float32 (M, K),
  ``selector_auxiliary_indices.bin`` int32 (only when the selector has any)
  and ``snapshot.json``.

Binary files are the raw little-endian arrays with no header.  The recorder
takes host arrays; the caller synchronizes and copies off the device, as the
reference does before each write (data_recorder.cu:14-34, 219).
"""
from __future__ import annotations

import json
import os
from pathlib import Path
from typing import Any

import numpy as np

LAYOUTS = {
    "variables": "row-major (N,D)",
    "objectives": "objective-major (M,N)",
    "constraints": "(N)",
    "reference_directions": "objective-major (M,K)",
    "bounds": "separate lower/upper arrays (D)",
}


def _write(path: Path, values: Any, dtype: Any) -> None:
    data = np.ascontiguousarray(np.asarray(values, dtype=dtype))
    with open(path, "wb") as out:
        out.write(data.astype(data.dtype.newbyteorder("<"), copy=False).tobytes())


def _write_json(path: Path, value: dict[str, Any]) -> None:
    # two-space indentation, "key": value, raw UTF-8 (data_recorder.cu:157-185, 260-273)
    with open(path, "w", encoding="utf-8", newline="\n") as out:
        out.write(json.dumps(value, indent=2, ensure_ascii=False) + "\n")


def generation_directory(root: str | os.PathLike[str], generation: int) -> Path:
    """data_recorder.cu:68-76: ``generation_`` and six zero-padded digits."""
    return Path(root) / f"generation_{generation:06d}"


class SnapshotRecorder:
    """Writes the run's metadata once and one directory per saved generation.

    ``should_save`` follows data_recorder.cu:129-136: generation 0, the last
    generation and every multiple of ``interval``."""

    def __init__(self, directory: str | os.PathLike[str], interval: int = 1) -> None:
        interval = int(interval)
        if interval <= 0:
            raise ValueError("saveData generation interval must be positive")
        if not str(directory):
            raise ValueError("saveData output directory cannot be empty")
        self.root = Path(directory)
        self.interval = interval
        self._ready = False

    def should_save(self, generation: int, max_generations: int) -> bool:
        return generation == 0 or generation == max_generations or generation % self.interval == 0

    def next_save(self, generation: int, max_generations: int) -> int:
        """First generation after ``generation`` that ``should_save`` accepts."""
        return min((generation // self.interval + 1) * self.interval, max_generations)

    def initialize(self, *, algorithm_name: str | None, problem_name: str | None, population_size: int,
                   dimension: int, objective_count: int, constraint_count: int, max_generations: int,
                   random_seed: int, lower_bounds: Any, upper_bounds: Any,
                   initial_directions: np.ndarray | None) -> None:
        """data_recorder.cu:138-209."""
        self.root.mkdir(parents=True, exist_ok=True)
        _write(self.root / "lower_bounds.bin", lower_bounds, np.float32)
        _write(self.root / "upper_bounds.bin", upper_bounds, np.float32)
        k = 0
        if initial_directions is not None and initial_directions.size > 0:
            k = int(initial_directions.shape[1])
            _write(self.root / "initial_reference_directions.bin", initial_directions, np.float32)
        _write_json(self.root / "metadata.json", {
            "format_version": 1,
            "scalar_type": "float32",
            "index_type": "int32",
            "algorithm_name": algorithm_name or None,
            "problem_name": problem_name or None,
            "population_size": int(population_size),
            "dimension": int(dimension),
            "objective_count": int(objective_count),
            "constraint_count": int(constraint_count),
            "max_generations": int(max_generations),
            "random_seed": int(random_seed),
            "generation_interval": self.interval,
            "reference_direction_count": k,
            "layouts": dict(LAYOUTS),
        })
        self._ready = True

    def save(self, generation: int, variables: np.ndarray, objectives: np.ndarray, constraints: np.ndarray,
             directions: np.ndarray | None, auxiliary: np.ndarray | None, active_count: int) -> Path:
        """data_recorder.cu:211-290.  ``objectives`` is (M, N), objective-major."""
        if not self._ready:
            raise RuntimeError("DataRecorder must be initialized before saving snapshots")
        n, d = variables.shape
        m = objectives.shape[0]
        if objectives.shape != (m, n) or constraints.size != n:
            raise ValueError("snapshot arrays disagree on the population size")
        out = generation_directory(self.root, generation)
        out.mkdir(parents=True, exist_ok=True)
        _write(out / "variables.bin", variables, np.float32)
        _write(out / "objectives.bin", objectives, np.float32)
        _write(out / "constraints.bin", constraints, np.float32)
        k = 0
        if directions is not None and directions.size > 0:
            k = int(directions.shape[1])
            _write(out / "reference_directions.bin", directions, np.float32)
        aux = 0
        if auxiliary is not None and auxiliary.size > 0:
            aux = int(auxiliary.size)
            _write(out / "selector_auxiliary_indices.bin", auxiliary, np.int32)
        _write_json(out / "snapshot.json", {
            "generation": int(generation),
            "population_size": int(n),
            "dimension": int(d),
            "objective_count": int(m),
            "active_count": int(active_count),
            "reference_direction_count": k,
            "auxiliary_index_count": aux,
        })
        return out


def load_snapshot(root: str | os.PathLike[str], generation: int) -> dict[str, Any]:
    """Reads one saved generation back: the arrays in their recorded shapes
    and the ``snapshot.json`` fields, using the run's ``metadata.json``."""
    root = Path(root)
    with open(root / "metadata.json", encoding="utf-8") as fh:
        meta = json.load(fh)
    gdir = generation_directory(root, generation)
    with open(gdir / "snapshot.json", encoding="utf-8") as fh:
        snap = json.load(fh)
    n, d, m = snap["population_size"], snap["dimension"], snap["objective_count"]
    k = snap["reference_direction_count"]
    out: dict[str, Any] = dict(snap)
    out["metadata"] = meta
    out["variables"] = np.fromfile(gdir / "variables.bin", dtype="<f4").reshape(n, d)
    out["objectives"] = np.fromfile(gdir / "objectives.bin", dtype="<f4").reshape(m, n)
    out["constraints"] = np.fromfile(gdir / "constraints.bin", dtype="<f4").reshape(n)
    out["reference_directions"] = (np.fromfile(gdir / "reference_directions.bin", dtype="<f4").reshape(m, k)
                                   if k > 0 else None)
    out["auxiliary_indices"] = (np.fromfile(gdir / "selector_auxiliary_indices.bin", dtype="<i4")
                                if snap["auxiliary_index_count"] > 0 else None)
    return out
