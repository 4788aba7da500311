"""MI355X-native MOEA engine with the API of the reference's ``cuda_moea``
package (python/cuda_moea/__init__.py): NSGA-III and RVEA generations run as
hand-written gfx950 HIP kernels in libmoea_hip.so."""
import torch  # noqa: F401  (device memory and streams)

from .algorithm import Algorithm, AlgorithmBuilder, NSGA3, Population, RVEA, Result
from .batch import BatchNSGA3
from .serve import Server
from .record import SnapshotRecorder, load_snapshot
from .native import NativeProblem
from . import problem as _problem
from .problem import *  # noqa: F401,F403
from ._specs import (
    AdaptiveRVEADirections, CudaConfig, DasDennisDirections,
    NSGA3EnvironmentSelector, NoMutation, PolynomialMutation,
    PythonCrossover, PythonEnvironmentSelector, PythonMating, PythonMutation,
    PythonReferenceDirections, RandomMating, RVEAEnvironmentSelector, SBX,
    TournamentMating, UserDefinedDirections,
)

__version__ = "0.2.0"

__all__ = [
    "BatchNSGA3", "Server", "SnapshotRecorder", "load_snapshot", "NativeProblem",
    "Algorithm", "AlgorithmBuilder", "NSGA3", "RVEA", "Population", "Result",
    "CudaConfig", "SBX", "PolynomialMutation", "NoMutation",
    "RandomMating", "TournamentMating", "DasDennisDirections",
    "AdaptiveRVEADirections", "UserDefinedDirections",
    "NSGA3EnvironmentSelector", "RVEAEnvironmentSelector",
    "PythonProblem", "PythonMating", "PythonCrossover", "PythonMutation",
    "PythonReferenceDirections", "PythonEnvironmentSelector",
] + _problem.__all__
