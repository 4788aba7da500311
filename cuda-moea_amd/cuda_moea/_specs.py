"""Strategy specifications, field-for-field those of the reference
(python/cuda_moea/_specs.py:14-153), so user code constructs them unchanged."""
from __future__ import annotations

from dataclasses import asdict, dataclass
from typing import Any, Sequence


class _Spec:
    def _spec(self) -> dict[str, Any]:
        data = asdict(self) if hasattr(self, "__dataclass_fields__") else {}
        data["type"] = type(self).__name__
        return data


@dataclass
class CudaConfig:
    device_id: int = 0
    evaluation_pool_ratio: float = 0.25
    execution_pool_ratio: float = 0.75
    memory_pool_policy: int = 1
    seed: int = 2887
    enable_warmup: bool = True


@dataclass
class SBX(_Spec):
    eta_initial: float = 30.0
    eta_final: float = 30.0
    probability: float = 1.0
    variable_copy_probability: float = 0.0

    def validate(self) -> None:
        # sbx_crossover.cu:86-99
        if (self.eta_initial <= 0 or self.eta_final <= 0 or not 0 <= self.probability <= 1
                or not 0 <= self.variable_copy_probability <= 1):
            raise ValueError("Invalid SBX configuration")


@dataclass
class PolynomialMutation(_Spec):
    eta_initial: float = 20.0
    eta_final: float = 20.0
    probability: float = 1.0

    def validate(self) -> None:
        # polynomial_mutation.cu:64-75
        if self.eta_initial <= 0 or self.eta_final <= 0 or not 0 <= self.probability <= 1:
            raise ValueError("Invalid polynomial-mutation configuration")


@dataclass
class NoMutation(_Spec):
    pass


@dataclass
class RandomMating(_Spec):
    pass


@dataclass
class TournamentMating(_Spec):
    pass


@dataclass
class DasDennisDirections(_Spec):
    partitions: int = 0


@dataclass
class AdaptiveRVEADirections(_Spec):
    frequency: float = 0.1


@dataclass
class UserDefinedDirections(_Spec):
    values: Any = None
    normalize: bool = True


@dataclass
class NSGA3EnvironmentSelector(_Spec):
    sparse_ratio: float = 0.5
    cv_bins: int = 0
    cv_clip_upper: float = 0.0
    cv_log_alpha: float = 0.0
    feasibility_epsilon: float = 0.0


@dataclass
class RVEAEnvironmentSelector(_Spec):
    alpha: float = 2.0


class PythonProblem(_Spec):
    """User problem evaluated in Python on device tensors
    (reference python/cuda_moea/_specs.py:86-107, adapters.cpp:150-190).

    ``evaluate(variables, context)`` receives a (population_size, dimension)
    float32 CUDA tensor and returns objectives (population_size, objectives),
    or a tuple / dict with ``objectives`` and optional ``constraints``."""

    def __init__(self, dimension: int, objectives: int,
                 lower_bounds: float | Sequence[float] = 0.0,
                 upper_bounds: float | Sequence[float] = 1.0,
                 constraints: int = 0, name: str | None = None) -> None:
        self.dimension = dimension
        self.objectives = objectives
        self.lower_bounds = _expand(lower_bounds, dimension)
        self.upper_bounds = _expand(upper_bounds, dimension)
        self.constraints = constraints
        self.name = name or type(self).__name__

    def _spec(self) -> dict[str, Any]:
        return {"type": "PythonProblem", "object": self, "dimension": self.dimension,
                "objectives": self.objectives, "lower_bounds": self.lower_bounds,
                "upper_bounds": self.upper_bounds, "constraints": self.constraints,
                "name": self.name}

    def evaluate(self, variables: Any, context: dict[str, Any]) -> Any:
        raise NotImplementedError


class _PythonStrategy(_Spec):
    kind = "PythonStrategy"

    def _spec(self) -> dict[str, Any]:
        return {"type": type(self).kind, "object": self}


class PythonMating(_PythonStrategy):
    kind = "PythonMating"

    def select(self, parents, state, context):
        raise NotImplementedError


class PythonCrossover(_PythonStrategy):
    kind = "PythonCrossover"

    def apply(self, parents, parent_indices, context):
        raise NotImplementedError


class PythonMutation(_PythonStrategy):
    kind = "PythonMutation"

    def mutate(self, offspring, context):
        raise NotImplementedError


class PythonReferenceDirections(_PythonStrategy):
    kind = "PythonReferenceDirections"

    def initialize(self, requested_count, objective_count, context):
        raise NotImplementedError


class PythonEnvironmentSelector(_PythonStrategy):
    kind = "PythonEnvironmentSelector"

    def initialize(self, info, references, context):
        pass

    def prepare(self, population, context):
        return None

    def select(self, parents, offspring, context):
        raise NotImplementedError


def _expand(value: float | Sequence[float], size: int) -> list[float]:
    if isinstance(value, (int, float)):
        return [float(value)] * size
    result = [float(v) for v in value]
    if len(result) != size:
        raise ValueError(f"expected {size} bounds, got {len(result)}")
    return result


def specification(value: Any) -> dict[str, Any]:
    if not hasattr(value, "_spec"):
        raise TypeError(f"{value!r} is not a cuda_moea strategy")
    return value._spec()
