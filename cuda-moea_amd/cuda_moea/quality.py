"""Quality of finished runs: analytic DTLZ Pareto fronts, IGD and exact
hypervolume, and the NSGA-III direction set of the reference's DTLZ quality
protocol.

The front samplers draw the same torch CPU random streams, in the same
order, as the reference's evaluation code (tests/evaluation/pareto_front.py:
17-198), so a front of a given size and seed is the same point set.  IGD follows
the reference's offline analysis (tests/benchmark/DTLZ/analyze_results.py:
88-110): finite rows, exact duplicates removed, dominated rows removed, then
the mean distance from each front point to its nearest kept row.  The
hypervolume is exact for two and three objectives (a sweep over the last
objective of exact two-dimensional areas); the reference estimates three
objectives with Sobol samples (tests/evaluation/metrics/hypervolume.py:37-70).
"""
from __future__ import annotations

import math

import numpy as np

# the reference's fixed front seed and point counts (TEST_PLAN.md 6.1 of the
# reference's benchmark plan; DTLZ4's ranking changed at 20,000 points, so its
# published IGD uses 50,000)
FRONT_SEED = 2887
FRONT_POINTS = {"DTLZ4": 50000}


def _canonical(name: str) -> str:
    return "".join(c for c in name.lower() if c.isalnum())


def _torch():
    import torch
    return torch


def _simplex(count, m, scale, seed):
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(seed)
    u = torch.rand((count, m), generator=g, dtype=torch.float32)
    e = -torch.log(u.clamp_min(torch.finfo(torch.float32).tiny))
    return scale * e / e.sum(dim=1, keepdim=True)


def _sphere(count, m, seed):
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(seed)
    p = torch.rand((count, m), generator=g, dtype=torch.float32)
    return p / p.norm(dim=1, keepdim=True).clamp_min(torch.finfo(torch.float32).eps)


def _from_angles(theta):
    """DTLZ2's position map of unit angles (theta in [0, pi/2])."""
    torch = _torch()
    n, k = theta.shape
    m = k + 1
    out = torch.ones((n, m), dtype=theta.dtype)
    for j in range(m):
        cos_count = m - 1 - j
        if cos_count:
            out[:, j] *= torch.cos(theta[:, :cos_count]).prod(dim=1)
        if j:
            out[:, j] *= torch.sin(theta[:, m - 1 - j])
    return out


def pareto_front_points(problem: str, objectives: int = 3, count: int = 20000, seed: int = FRONT_SEED) -> np.ndarray:
    """`count` points of the analytic Pareto front of an unconstrained DTLZ
    problem (float64 copy of the float32 samples), as
    tests/evaluation/pareto_front.py:142-175 of the reference draws them:

    - DTLZ1: the simplex sum f = 0.5 (exponential spacings);
    - DTLZ2/3/4: the positive unit sphere (normalised uniform cube points);
    - DTLZ5/6: the degenerate curve theta_0 uniform in [0, pi/2), the other
      angles pi/4;
    - DTLZ7: the disconnected front, free coordinates in the two
      intervals [0, 0.2514], [0.6316, 0.8594], the last 2M - sum f(1 + sin 3 pi f);
    - ConvexDTLZ2: the sphere with f_i^4 (i < M-1) and f_{M-1}^2.
    """
    torch = _torch()
    if objectives < 2 or count <= 0:
        raise ValueError("need at least two objectives and a positive point count")
    p = _canonical(problem)
    m = objectives
    if p == "dtlz1":
        out = _simplex(count, m, 0.5, seed)
    elif p in ("dtlz2", "dtlz3", "dtlz4"):
        out = _sphere(count, m, seed)
    elif p in ("dtlz5", "dtlz6"):
        g = torch.Generator(device="cpu").manual_seed(seed)
        theta = torch.full((count, m - 1), math.pi / 4.0, dtype=torch.float32)
        theta[:, 0] = torch.rand(count, generator=g, dtype=torch.float32) * (math.pi / 2.0)
        out = _from_angles(theta)
    elif p == "dtlz7":
        g = torch.Generator(device="cpu").manual_seed(seed)
        intervals = torch.tensor(((0.0, 0.251411836), (0.631626531, 0.859400856)), dtype=torch.float32)
        region = torch.randint(0, 2, (count, m - 1), generator=g)
        unit = torch.rand((count, m - 1), generator=g, dtype=torch.float32)
        free = intervals[region, 0] + unit * (intervals[region, 1] - intervals[region, 0])
        last = 2.0 * m - (free * (1.0 + torch.sin(3.0 * math.pi * free))).sum(dim=1, keepdim=True)
        out = torch.cat((free, last), dim=1)
    elif p == "convexdtlz2":
        out = _sphere(count, m, seed).clone()
        out[:, :-1] = out[:, :-1].pow(4)
        out[:, -1] = out[:, -1].pow(2)
    else:
        raise NotImplementedError(f"no analytic Pareto front for {problem!r}")
    return out.numpy().astype(np.float64)


def nondominated(points: np.ndarray) -> np.ndarray:
    """Mask of the rows no other row dominates (minimisation)."""
    pts = np.asarray(points, np.float64)
    n = len(pts)
    keep = np.ones(n, dtype=bool)
    for start in range(0, n, 1024):
        block = pts[start:start + 1024]
        le = (pts[:, None, :] <= block[None, :, :]).all(axis=2)
        lt = (pts[:, None, :] < block[None, :, :]).any(axis=2)
        keep[start:start + 1024] = ~(le & lt).any(axis=0)
    return keep


def clean_front(points) -> np.ndarray:
    """Finite rows, exact duplicates removed, dominated rows removed
    (analyze_results.py:88-100 of the reference)."""
    pts = np.asarray(points, np.float64)
    pts = pts[np.isfinite(pts).all(axis=1)]
    pts = np.unique(pts, axis=0)
    if not len(pts):
        raise ValueError("no finite objective rows")
    return pts[nondominated(pts)]


def igd(front, reference) -> float:
    """Mean Euclidean distance from each reference point to its nearest
    point of the cleaned `front`."""
    from scipy.spatial import cKDTree
    f = clean_front(front)
    dist = cKDTree(f).query(np.asarray(reference, np.float64), k=1)[0]
    return float(np.mean(dist))


def _hv2(pts: np.ndarray, ref: np.ndarray) -> float:
    """Area dominated by 2-D points (all strictly inside the reference box)."""
    if not len(pts):
        return 0.0
    order = np.lexsort((pts[:, 1], pts[:, 0]))
    x, y = pts[order, 0], np.minimum.accumulate(pts[order, 1])
    nxt = np.append(x[1:], ref[0])
    return float(np.sum((nxt - x) * (ref[1] - y)))


def hypervolume(points, reference_point, maximize: bool = False) -> float:
    """Exact volume dominated by `points` (minimisation, or maximisation with
    `maximize`) and bounded by `reference_point`, for two or three objectives
    (the exact method of tests/evaluation/metrics.py's hypervolume).  Rows not
    strictly better than the reference point in every objective contribute
    nothing."""
    pts = np.asarray(points, np.float64)
    ref = np.asarray(reference_point, np.float64)
    if maximize:
        pts, ref = -pts, -ref
    if pts.ndim != 2 or ref.shape != (pts.shape[1],) or not np.isfinite(ref).all():
        raise ValueError("reference_point must hold one finite value per objective")
    pts = pts[np.isfinite(pts).all(axis=1)]
    pts = pts[(pts < ref).all(axis=1)]
    if not len(pts):
        return 0.0
    m = pts.shape[1]
    if m == 2:
        return _hv2(pts, ref)
    if m != 3:
        raise ValueError("exact hypervolume is implemented for two and three objectives")
    pts = pts[np.argsort(pts[:, 2], kind="stable")]
    z = np.append(pts[:, 2], ref[2])
    vol = 0.0
    for i in range(len(pts)):
        depth = z[i + 1] - z[i]
        if depth > 0.0:
            vol += depth * _hv2(pts[:i + 1, :2], ref[:2])
    return float(vol)


def hv_reference_point(front: np.ndarray) -> np.ndarray:
    """Nadir + 10 % of the front's range (tests/evaluation/evaluate_dtlz.py:
    62-70 of the reference)."""
    lo, hi = front.min(axis=0), front.max(axis=0)
    return hi + 0.1 * np.maximum(hi - lo, np.finfo(np.float32).eps)


def uniform_directions(n: int, m: int) -> np.ndarray:
    """The NSGA-III direction set of the reference's DTLZ quality runs
    (tests/benchmark/adapters/shared.py:27-43 there): EvoX 1.3.0's
    ``uniform_sampling(n, m)``, which restates PlatEMO's UniformPoint in its
    normal-boundary-intersection form (Das and Dennis 1998).  H1 is the
    first H with C(H + m, m - 1) > n.  The C(H1 + m - 1, m - 1) points with
    coordinates k / H1 (k summing to H1) are enumerated by stars and bars:
    each (m - 1)-subset of bar positions, in lexicographic order, gives the
    cumulative sums s_j = c_j - j (j = 1..m-1), and w = diff(0, s, H1) / H1.
    When H1 < m an inner layer with H2 partitions, halved and shifted by
    1 / (2m), is appended.  Coordinates are clamped below at 1e-6.  float32."""
    from itertools import combinations

    def layer(h):
        c = np.array(list(combinations(range(1, h + m), m - 1)), dtype=np.int64) - np.arange(m - 1) - 1
        rows = len(c)
        upper = np.concatenate([c, np.full((rows, 1), h, np.int64)], axis=1)
        lower = np.concatenate([np.zeros((rows, 1), np.int64), c], axis=1)
        return (upper - lower).astype(np.float32) / np.float32(h)

    h1 = 1
    while math.comb(h1 + m, m - 1) <= n:
        h1 += 1
    w = layer(h1)
    if h1 < m:
        h2 = 0
        while math.comb(h1 + m - 1, m - 1) + math.comb(h2 + m, m - 1) <= n:
            h2 += 1
        if h2 > 0:
            w = np.concatenate([w, layer(h2) / np.float32(2) + np.float32(1.0 / (2 * m))], axis=0)
    return np.maximum(w, np.float32(1e-6)).astype(np.float32)
