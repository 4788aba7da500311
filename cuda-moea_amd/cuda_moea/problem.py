"""Problems.

Built-in HIP objectives (evaluated by libmoea_hip.so inside the generation
loop): :class:`QuadraticCenters` and :class:`SeededMLP`.  The DTLZ family of
the reference (python/cuda_moea/problem.py:9-47, kernels in
NS/problem/dtlz/dtlz_kernels.cu) is provided as device-tensor Python problems.
"""
from __future__ import annotations

import math
from typing import Any

import torch

from ._specs import PythonProblem


class QuadraticCenters:
    """f_c(x) = ||x - c_c||^2 for m centers c_c in [0, 1]^d.

    Centers come from ``torch.Generator('cpu').manual_seed(seed)`` as
    ``torch.rand(m, d)`` unless given explicitly."""

    def __init__(self, dimension: int, objectives: int = 3, seed: int = 0,
                 centers: Any = None, lower_bounds: float = 0.0,
                 upper_bounds: float = 1.0) -> None:
        self.dimension = int(dimension)
        self.objectives = int(objectives)
        if centers is None:
            g = torch.Generator(device="cpu").manual_seed(int(seed))
            centers = torch.rand(self.objectives, self.dimension, generator=g)
        self.centers = torch.as_tensor(centers, dtype=torch.float32).reshape(
            self.objectives, self.dimension).contiguous()
        self.lower_bounds = [float(lower_bounds)] * self.dimension
        self.upper_bounds = [float(upper_bounds)] * self.dimension
        self.constraints = 0
        self.name = "QuadraticCenters"

    def _spec(self) -> dict[str, Any]:
        return {"type": "QuadraticCenters", "object": self}


class SeededMLP:
    """f_c(x) = y_c^2 + 0.01 ||x - 0.5||^2 with y = W2 tanh(W1 x), W1 ~ N(0, 1/d)
    (hidden x d), W2 ~ N(0, 1/hidden) (objectives x hidden), seeded from
    ``torch.Generator(device).manual_seed(seed)`` (CPU by default; a GPU
    generator draws other weights of the same distribution, in place)."""

    def __init__(self, dimension: int, objectives: int = 5, hidden: int = 256,
                 seed: int = 0, device: Any = "cpu") -> None:
        self.dimension = int(dimension)
        self.objectives = int(objectives)
        self.hidden = int(hidden)
        g = torch.Generator(device=device).manual_seed(int(seed))
        self.w1 = (torch.randn(self.hidden, self.dimension, generator=g, device=device)
                   / math.sqrt(self.dimension)).contiguous()
        self.w2 = (torch.randn(self.objectives, self.hidden, generator=g, device=device)
                   / math.sqrt(self.hidden)).contiguous()
        self.lower_bounds = [0.0] * self.dimension
        self.upper_bounds = [1.0] * self.dimension
        self.constraints = 0
        self.name = "SeededMLP"

    def _spec(self) -> dict[str, Any]:
        return {"type": "SeededMLP", "object": self}


class DTLZProblem(PythonProblem):
    """DTLZ1-7 and ConvexDTLZ2 (Deb, Thiele, Laumanns, Zitzler 2005), the
    constrained C1/C2/C3 variants and CSDP, with the reference's constructor
    (problem.py:9-47).  The algorithm evaluates them with the built-in
    kernels; ``evaluate`` is the FP64 torch form of the unconstrained ones,
    used where the kernels do not apply (one objective, or more objectives
    than variables)."""

    _kind = "DTLZ2"
    _constrained = ("C1DTLZ1", "C1DTLZ3", "C2DTLZ2", "C2ConvexDTLZ2", "C3DTLZ1", "C3DTLZ4", "CSDP")

    def __init__(self, dimension: int = 12, objectives: int = 3,
                 constraint_activation_ratio: float = 0.0) -> None:
        # dtlz_problem.cu:63-77: finite, clamped to [0, 1]; 0 keeps constraints on from the start
        if not math.isfinite(constraint_activation_ratio):
            raise ValueError("constraint_activation_ratio must be finite")
        self.constraint_activation_ratio = min(max(float(constraint_activation_ratio), 0.0), 1.0)
        constrained = self._kind in self._constrained
        super().__init__(dimension, objectives, 0.0, 1.0, 1 if constrained else 0, self._kind)

    def evaluate(self, variables: Any, context: dict[str, Any]) -> Any:
        x = variables.double()
        m = self.objectives
        xm, xk = x[:, : m - 1], x[:, m - 1:]
        kind = self._kind
        if kind in self._constrained:
            raise NotImplementedError(f"{kind} runs on the built-in kernels only (2 <= objectives <= dimension)")
        if kind in ("DTLZ1", "DTLZ3"):
            g = 100.0 * (xk.shape[1] + ((xk - 0.5) ** 2 - torch.cos(20 * math.pi * (xk - 0.5))).sum(1))
        elif kind in ("DTLZ2", "DTLZ4", "ConvexDTLZ2", "DTLZ5"):
            g = ((xk - 0.5) ** 2).sum(1)
        elif kind == "DTLZ6":
            g = (xk ** 0.1).sum(1)
        elif kind == "DTLZ7":
            g = 1.0 + 9.0 / xk.shape[1] * xk.sum(1)
        else:
            raise NotImplementedError(kind)
        n = x.shape[0]
        f = torch.empty(n, m, dtype=torch.float64, device=x.device)
        if kind == "DTLZ1":
            for i in range(m):
                v = 0.5 * (1 + g)
                v = v * torch.prod(xm[:, : m - 1 - i], dim=1) if m - 1 - i > 0 else v
                if i > 0:
                    v = v * (1 - xm[:, m - 1 - i])
                f[:, i] = v
        elif kind == "DTLZ7":
            f[:, : m - 1] = xm
            h = m - (xm / (1 + g[:, None]) * (1 + torch.sin(3 * math.pi * xm))).sum(1)
            f[:, m - 1] = (1 + g) * h
        else:
            theta = xm.clone()
            if kind == "DTLZ4":
                theta = theta ** 100
            if kind in ("DTLZ5", "DTLZ6"):
                t = 1.0 / (2 * (1 + g[:, None])) * (1 + 2 * g[:, None] * xm)
                theta = torch.cat([xm[:, :1], t[:, 1:]], dim=1)
            ang = theta * (math.pi / 2)
            for i in range(m):
                v = 1 + g
                if m - 1 - i > 0:
                    v = v * torch.prod(torch.cos(ang[:, : m - 1 - i]), dim=1)
                if i > 0:
                    v = v * torch.sin(ang[:, m - 1 - i])
                f[:, i] = v
            if kind == "ConvexDTLZ2":
                f[:, : m - 1] = f[:, : m - 1] ** 4
                f[:, m - 1] = f[:, m - 1] ** 2
        return f.float()


def _problem_class(name: str):
    class BuiltinProblem(DTLZProblem):
        _kind = name

        def __init__(self, dimension: int = 12, objectives: int = 3,
                     constraint_activation_ratio: float = 0.0) -> None:
            super().__init__(dimension, objectives, constraint_activation_ratio)

    BuiltinProblem.__name__ = name
    BuiltinProblem.__qualname__ = name
    return BuiltinProblem


DTLZ1 = _problem_class("DTLZ1")
DTLZ2 = _problem_class("DTLZ2")
DTLZ3 = _problem_class("DTLZ3")
DTLZ4 = _problem_class("DTLZ4")
DTLZ5 = _problem_class("DTLZ5")
DTLZ6 = _problem_class("DTLZ6")
DTLZ7 = _problem_class("DTLZ7")
ConvexDTLZ2 = _problem_class("ConvexDTLZ2")
C1DTLZ1 = _problem_class("C1DTLZ1")
C1DTLZ3 = _problem_class("C1DTLZ3")
C2DTLZ2 = _problem_class("C2DTLZ2")
C2ConvexDTLZ2 = _problem_class("C2ConvexDTLZ2")
C3DTLZ1 = _problem_class("C3DTLZ1")
C3DTLZ4 = _problem_class("C3DTLZ4")
CSDP = _problem_class("CSDP")

__all__ = ["PythonProblem", "QuadraticCenters", "SeededMLP", "DTLZProblem", "DTLZ1", "DTLZ2",
           "DTLZ3", "DTLZ4", "DTLZ5", "DTLZ6", "DTLZ7", "ConvexDTLZ2", "C1DTLZ1", "C1DTLZ3",
           "C2DTLZ2", "C2ConvexDTLZ2", "C3DTLZ1", "C3DTLZ4", "CSDP"]
