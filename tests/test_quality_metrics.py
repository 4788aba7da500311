"""Quality indicators (cuda_moea.quality) on the CPU: exact hypervolume
against brute force over the coordinate grid and closed forms, IGD identities,
front cleaning, the analytic front samplers against the fronts' defining
equations, the NSGA-III uniform direction set, and the published-IGD fixture
against the reference's own summary medians."""
import csv
import itertools
import math
from pathlib import Path

import numpy as np
import pytest

from cuda_moea import quality as Q

GOLDEN = Path(__file__).resolve().parent / "golden" / "dtlz_quality_reference.csv"


def _hv_brute(points, ref):
    """Union volume of the boxes [p, ref] by cell enumeration over the grid of
    all coordinates: a cell is covered when its lower corner is weakly
    dominated by some point."""
    pts = np.asarray(points, np.float64)
    pts = pts[(pts < ref).all(axis=1)]
    if not len(pts):
        return 0.0
    axes = [np.unique(np.append(pts[:, j], ref[j])) for j in range(pts.shape[1])]
    vol = 0.0
    for cell in itertools.product(*[range(len(a) - 1) for a in axes]):
        lo = np.array([axes[j][c] for j, c in enumerate(cell)])
        hi = np.array([axes[j][c + 1] for j, c in enumerate(cell)])
        if (pts <= lo).all(axis=1).any():
            vol += float(np.prod(hi - lo))
    return vol


@pytest.mark.parametrize("m", [2, 3])
@pytest.mark.parametrize("seed", range(6))
def test_hypervolume_matches_brute_force(m, seed):
    rng = np.random.default_rng(seed)
    n = 9 if m == 3 else 25
    pts = rng.random((n, m))
    pts[1] = pts[0]                                  # duplicate
    pts[2] = pts[0] + 0.05                           # dominated
    pts[3, 0] = 1.3                                  # outside the box
    pts = np.round(pts, 2)                           # ties between coordinates
    ref = np.full(m, 1.1)
    assert Q.hypervolume(pts, ref) == pytest.approx(_hv_brute(pts, ref), rel=1e-12, abs=1e-15)


def test_hypervolume_closed_forms():
    assert Q.hypervolume([[0.2, 0.3, 0.4]], [1.0, 1.0, 1.0]) == pytest.approx(0.8 * 0.7 * 0.6)
    assert Q.hypervolume([[0.2, 0.5]], [1.0, 1.0]) == pytest.approx(0.4)
    # a 2-D staircase (0, 1), (0.5, 0.5), (1, 0) against (2, 2): 2 + 0.75 + 0.75 + ... by areas
    assert Q.hypervolume([[0, 1], [0.5, 0.5], [1, 0]], [2, 2]) == pytest.approx(2 * 1 + 1.5 * 0.5 + 1 * 0.5)
    assert Q.hypervolume([[1.0, 1.0, 1.0]], [1.0, 1.0, 1.0]) == 0.0
    # the positive unit-sphere octant dominates 1 - pi/6 of the unit cube; a
    # sample approaches it from below as the sample grows
    small = Q.hypervolume(Q.pareto_front_points("DTLZ2", 3, 500), np.ones(3))
    large = Q.hypervolume(Q.pareto_front_points("DTLZ2", 3, 4000), np.ones(3))
    exact = 1.0 - math.pi / 6.0
    assert small < large < exact and large > 0.96 * exact
    # the simplex sum f = 0.5 leaves 1 - 0.5^3 / 6 of the unit cube
    simplex = Q.hypervolume(Q.pareto_front_points("DTLZ1", 3, 4000), np.ones(3))
    assert 0.99 * (1 - 0.5 ** 3 / 6) < simplex < 1 - 0.5 ** 3 / 6
    with pytest.raises(ValueError):
        Q.hypervolume(np.zeros((2, 4)), np.ones(4))


def test_hypervolume_reference_known_answers():
    """The known answers of the reference's hypervolume tests
    (tests/python/test_evaluation_metrics.py:26-58): one point, a union with a
    dominated point, points outside the reference ignored, maximisation, and an
    exact three-objective box."""
    assert Q.hypervolume([[1.0, 2.0]], [3.0, 4.0]) == pytest.approx(4.0)
    assert Q.hypervolume([[1.0, 2.0], [2.0, 1.0], [2.5, 2.5]], [3.0, 3.0]) == pytest.approx(3.0)
    assert Q.hypervolume([[1.0, 1.0], [4.0, 0.0]], [3.0, 3.0]) == pytest.approx(4.0)
    assert Q.hypervolume([[2.0, 1.0], [1.0, 2.0]], [0.0, 0.0], maximize=True) == pytest.approx(3.0)
    assert Q.hypervolume([[1.0, 1.0, 1.0]], [3.0, 4.0, 5.0]) == pytest.approx(24.0)


def test_igd_identities():
    ref = Q.pareto_front_points("DTLZ2", 3, 2000)
    assert Q.igd(ref, ref) == 0.0
    # duplicates, dominated and non-finite rows do not change the value
    front = Q.pareto_front_points("DTLZ2", 3, 300, seed=5)
    noisy = np.concatenate([front, front[:10], front[:10] * 1.5, np.full((3, 3), np.nan)])
    assert Q.igd(noisy, ref) == Q.igd(front, ref)
    # one point: the mean distance to it
    p = np.array([[0.5, 0.5, 0.5]])
    assert Q.igd(p, ref) == pytest.approx(np.linalg.norm(ref - p, axis=1).mean(), rel=1e-12)
    with pytest.raises(ValueError):
        Q.igd(np.full((2, 3), np.nan), ref)


def test_clean_front():
    pts = np.array([[1, 2], [1, 2], [2, 1], [2, 2], [0.5, 3], [np.inf, 0], [np.nan, 1]], float)
    got = Q.clean_front(pts)
    assert sorted(map(tuple, got)) == [(0.5, 3.0), (1.0, 2.0), (2.0, 1.0)]


def test_front_samplers_satisfy_their_fronts():
    n = 5000
    f1 = Q.pareto_front_points("DTLZ1", 3, n)
    assert np.allclose(f1.sum(1), 0.5, atol=1e-6) and (f1 >= 0).all()
    for p in ("DTLZ2", "DTLZ3", "DTLZ4"):
        s = Q.pareto_front_points(p, 3, n)
        assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-6) and (s >= 0).all()
    d5 = Q.pareto_front_points("DTLZ5", 3, n)
    assert np.allclose(d5[:, 0], d5[:, 1], atol=1e-6)                 # the degenerate curve f1 = f2
    assert np.allclose(np.linalg.norm(d5, axis=1), 1.0, atol=1e-6)
    assert np.array_equal(d5, Q.pareto_front_points("DTLZ6", 3, n))
    d7 = Q.pareto_front_points("DTLZ7", 3, n)
    free = d7[:, :2]
    in_a = (free >= 0) & (free <= 0.2514119)
    in_b = (free >= 0.6316265) & (free <= 0.8594009)
    assert (in_a | in_b).all()
    # g = 1 on the front: f3 = (1 + g) h, h = M - sum f_i / (1 + g) (1 + sin 3 pi f_i)
    h = 3 - (free / 2 * (1 + np.sin(3 * np.pi * free))).sum(1)
    assert np.allclose(d7[:, 2], 2 * h, atol=1e-5)
    c = Q.pareto_front_points("ConvexDTLZ2", 3, n)
    s = Q.pareto_front_points("DTLZ2", 3, n)
    assert np.allclose(c[:, :2], s[:, :2] ** 4, rtol=1e-6) and np.allclose(c[:, 2], s[:, 2] ** 2, rtol=1e-6)
    # deterministic in the seed
    assert np.array_equal(s, Q.pareto_front_points("DTLZ2", 3, n))
    assert not np.array_equal(s, Q.pareto_front_points("DTLZ2", 3, n, seed=1))
    with pytest.raises(NotImplementedError):
        Q.pareto_front_points("ZDT1", 2, 10)


def test_uniform_directions():
    w = Q.uniform_directions(1024, 3)
    assert w.shape == (math.comb(43 + 2, 2), 3) == (990, 3)
    assert w.dtype == np.float32 and (w >= 1e-6).all()
    k = np.rint(np.where(w <= 1e-6, 0, w) * 43)                        # coordinates are k / 43
    assert np.allclose(np.where(w <= 1e-6, 0, w), k / 43, atol=1e-7) and (k.sum(1) == 43).all()
    # the bar positions in lexicographic order: first (0, 0, 43), last (43, 0, 0)
    assert tuple(k[0]) == (0, 0, 43) and tuple(k[1]) == (0, 1, 42) and tuple(k[-1]) == (43, 0, 0)
    assert len({tuple(r) for r in k}) == 990
    # a second, inner layer when the outer one has fewer partitions than objectives
    w9 = Q.uniform_directions(9, 3)
    assert w9.shape == (9, 3)
    inner = w9[6:]
    assert np.allclose(np.sort(inner, axis=1), [[1 / 6, 1 / 6, 2 / 3]] * 3, atol=1e-7)
    assert Q.uniform_directions(10, 3).shape == (10, 3)                 # H = 3: C(5, 2) = 10


# the reference's summary medians (tests/benchmark/DTLZ/results/data/
# quality_summary.csv there), which the per-run fixture must reproduce
PUBLISHED_MEDIANS = {("DTLZ1", "nsga3"): 0.0057404759985065145, ("DTLZ2", "rvea"): 0.01699097149746534,
                     ("DTLZ5", "nsga3"): 0.0013404823678190062, ("DTLZ7", "rvea"): 0.030097557716765602,
                     ("ConvexDTLZ2", "nsga3"): 0.01619243150652637}


def test_published_fixture():
    with GOLDEN.open(newline="") as f:
        rows = list(csv.DictReader(f))
    assert len(rows) == 480
    for (p, a), med in PUBLISHED_MEDIANS.items():
        vals = [float(r["final_igd"]) for r in rows if r["problem"] == p and r["algorithm"] == a]
        assert sorted(int(r["seed"]) for r in rows if r["problem"] == p and r["algorithm"] == a) == list(range(30))
        assert np.median(vals) == pytest.approx(med, rel=1e-12)
    for r in rows:
        sel = r["igd_50000"] if r["pareto_reference_points"] == "50000" else r["igd_20000"]
        assert float(r["final_igd"]) == float(sel)
        assert int(r["actual_population"]) == (1024 if r["algorithm"] == "nsga3" else 990)
