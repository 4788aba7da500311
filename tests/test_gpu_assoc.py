"""Association (moea_associate): the cone-pruned path against every-direction
evaluation and the oracle, bit for bit, including the cases where pruning
bounds are tight or degenerate (points on directions, duplicate directions,
zero / tiny / huge / non-finite points, negative components)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _run(fn, w, method):
    m = fn.shape[0]
    d, r = K.associate(torch.from_numpy(fn).to(DEV), torch.ones(m, dtype=torch.float32, device=DEV),
                       torch.from_numpy(np.ascontiguousarray(w)).to(DEV), method)
    return d.cpu().numpy(), r.cpu().numpy()


def _check(fn, w):
    d1, r1 = _run(fn, w, 1)
    d2, r2 = _run(fn, w, 2)
    np.testing.assert_array_equal(r2, r1)
    np.testing.assert_array_equal(d2.view(np.uint32), d1.view(np.uint32))
    do, ro = O.associate(fn, w)
    np.testing.assert_array_equal(r1, ro)
    np.testing.assert_array_equal(d1.view(np.uint32), do.view(np.uint32))
    return r1


@pytest.mark.parametrize("n_req,m,npts,seed", [(32768, 3, 20000, 0), (4000, 2, 5000, 1), (5000, 5, 4000, 2),
                                               (3000, 8, 3000, 3), (2500, 10, 2000, 4)])
def test_pruned_equals_brute_force_random(n_req, m, npts, seed):
    rng = np.random.default_rng(seed)
    w = O.das_dennis(n_req, m)
    print(f"associate m={m} k={w.shape[1]} n={npts}", flush=True)
    fn = (rng.random((m, npts)) * rng.uniform(0.1, 3.0, (1, npts))).astype(np.float32)
    _check(fn, w)


def test_pruned_degenerate_points():
    rng = np.random.default_rng(7)
    m = 3
    w = O.das_dennis(8000, m)
    k = w.shape[1]
    cols = []
    # points exactly on directions, scaled
    idx = rng.integers(0, k, 300)
    cols.append(w[:, idx] * rng.uniform(0.5, 2.0, 300).astype(np.float32))
    # zero, tiny and huge points
    cols.append(np.zeros((m, 4), np.float32))
    cols.append((rng.random((m, 50)) * 1e-30).astype(np.float32))
    cols.append((rng.random((m, 50)) * 1e19).astype(np.float32))
    cols.append((rng.random((m, 20)) * 3e38).astype(np.float32))
    # non-finite points
    bad = rng.random((m, 30)).astype(np.float32)
    bad[0, :10] = np.nan
    bad[1, 10:20] = np.inf
    bad[2, 20:] = -np.inf
    cols.append(bad)
    # negative components and points on the simplex edges
    cols.append((rng.random((m, 200)) - 0.3).astype(np.float32))
    edge = rng.random((m, 100)).astype(np.float32)
    edge[rng.integers(0, m, 100), np.arange(100)] = 0.0
    cols.append(edge)
    # points half-way between two directions (neighbours and random pairs):
    # equal or nearly equal d2 in two blocks
    u = w / np.linalg.norm(w, axis=0)
    j = rng.integers(0, k - 1, 200)
    l = np.concatenate([j[:100] + 1, rng.integers(0, k, 100)])
    cols.append(((u[:, j] + u[:, l]) * rng.uniform(0.5, 2.0, 200)).astype(np.float32))
    fn = np.ascontiguousarray(np.concatenate(cols, axis=1))
    _check(fn, w)


def test_pruned_duplicate_directions_take_smallest_index():
    rng = np.random.default_rng(8)
    m = 3
    base = O.das_dennis(3000, m)
    w = np.ascontiguousarray(np.concatenate([base, base[:, ::-1], base], axis=1))
    fn = np.ascontiguousarray(np.concatenate([base[:, :500] * 1.5, rng.random((m, 1000)).astype(np.float32)],
                                             axis=1).astype(np.float32))
    r = _check(fn, w)
    assert (r[:500] < base.shape[1]).all()


def test_pruned_unnormalised_and_zero_directions():
    rng = np.random.default_rng(9)
    m = 4
    w = O.das_dennis(6000, m) * rng.uniform(0.25, 2.0, (1, 1)).astype(np.float32)
    w = np.ascontiguousarray(w * rng.uniform(0.9, 1.1, (1, w.shape[1])).astype(np.float32))
    w[:, 1000] = 0.0
    fn = np.ascontiguousarray((rng.random((m, 3000)) * rng.uniform(0.1, 3.0, (1, 3000))).astype(np.float32))
    print(f"associate m={m} k={w.shape[1]} scaled directions", flush=True)
    _check(fn, w)
