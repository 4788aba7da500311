"""Normalization of moea_nsga3_select against the oracle (orc_normalize, the
reference's legacy path with the hyperplane solve in FP64): the extreme point
of every axis (merged-row index) and the fallback flag bit for bit, the
intercepts within 1e-6 relative of the oracle's FP64 values; on hand-built
cases with known intercepts, degenerate extremes, NaN rows, M = 10 with equal
rows, and the full-size selection input (2N = 65536 rows s*u, K = 32640)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _f(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _sphere(rng, m, n):
    g = np.abs(rng.standard_normal((n, m)))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    return np.ascontiguousarray((rng.uniform(1.0, 2.0, (n, 1)) * u).T.astype(np.float32))


def _select(fm, seed=0):
    m, nm = fm.shape
    n = nm // 2
    cv = np.zeros(nm, np.float32)
    w = O.das_dennis(n, m)
    ideal = fm[:, :n].min(axis=1)
    x = np.zeros((nm, 2), np.float32)
    par = K.Population(_f(x[:n]), _f(fm[:, :n]), _f(cv[:n]))
    ofs = K.Population(_f(x[n:]), _f(fm[:, n:]), _f(cv[n:]))
    norm = {}
    _, _, st = K.nsga3_select(par, ofs, _f(w), _f(ideal.copy()), seed=seed, generation=0, normalization=norm)
    ua, ub = O.draw_niche(seed, 0, nm)
    onorm = {}
    _, _, ost, _ = O.nsga3_select(fm, cv, w, ideal.copy(), ua, ub, normalization=onorm)
    st = st.cpu().numpy()
    assert st[0] == ost[0] and st[1] == ost[1] and st[2] == ost[2]
    assert st[4] > 0, "case must normalize (n_rem > 0)"
    return st, ost, norm["extremes"].cpu().numpy(), norm["intercepts"].cpu().numpy(), onorm


def _check(st, ost, ext, icp, onorm):
    assert st[3] == ost[3]                                      # fallback flag
    np.testing.assert_array_equal(ext, onorm["extremes"])
    a64 = onorm["intercepts"]
    assert np.all(np.abs(icp - a64) <= 1e-6 * np.abs(a64)), (icp, a64)


def _with_axes(rng, m, n, scales):
    """A merged population of 2n rows on a sphere shell lifted by 0.5 plus, in
    its first columns, one row per axis with value scales[i] on axis i and 0
    elsewhere.  The axis rows set the ideal point to 0 and are the extreme
    points (every other row's achievement value is at least 1e6 x 0.5), so the
    extreme-point matrix is diag(scales) and the intercepts are 1 / scales."""
    fm = _sphere(rng, m, 2 * n) + 0.5
    for i, s in enumerate(scales):
        col = np.zeros(m, np.float32)
        col[i] = s
        fm[:, i] = col
    return fm


@pytest.mark.parametrize("m", [2, 3, 5, 8])
def test_identity_extremes_give_unit_intercepts(m):
    rng = np.random.default_rng(m)
    fm = _with_axes(rng, m, 300, [1.0] * m)
    print(f"normalization identity m={m}", flush=True)
    st, ost, ext, icp, on = _select(fm)
    _check(st, ost, ext, icp, on)
    assert st[3] == 0
    assert list(ext) == list(range(m))
    np.testing.assert_allclose(icp, np.ones(m), rtol=1e-12)


@pytest.mark.parametrize("m", [3, 4, 6])
def test_diagonal_extremes_give_reciprocal_intercepts(m):
    rng = np.random.default_rng(10 + m)
    scales = [1.5 + 0.5 * i for i in range(m)]
    fm = _with_axes(rng, m, 300, scales)
    print(f"normalization diagonal m={m}", flush=True)
    st, ost, ext, icp, on = _select(fm)
    _check(st, ost, ext, icp, on)
    assert st[3] == 0
    np.testing.assert_allclose(icp, 1.0 / np.array(scales), rtol=1e-6)


def test_duplicate_extreme_takes_the_fallback():
    """One row at the ideal point is the extreme of every axis (equal rows of
    the extreme-point matrix, all zero): the solve is singular and the
    intercepts come from the row maxima."""
    rng = np.random.default_rng(21)
    m, n = 3, 300
    fm = _sphere(rng, m, 2 * n) + 0.5
    fm[:, 7] = 0.0
    print("normalization duplicate extremes", flush=True)
    st, ost, ext, icp, on = _select(fm)
    _check(st, ost, ext, icp, on)
    assert st[3] == 1 and list(ext) == [7, 7, 7]


def test_nan_rows():
    rng = np.random.default_rng(22)
    m, n = 3, 400
    fm = _sphere(rng, m, 2 * n)
    fm[:, rng.random(2 * n) < 0.05] = np.nan
    fm[1, rng.random(2 * n) < 0.03] = np.nan
    print("normalization NaN rows", flush=True)
    st, ost, ext, icp, on = _select(fm)
    _check(st, ost, ext, icp, on)


def test_ten_objectives_equal_rows():
    rng = np.random.default_rng(23)
    m, n = 10, 300
    fm = _sphere(rng, m, 2 * n)
    fm[:, 5:40] = fm[:, 5:6]                                     # many equal rows
    print("normalization m=10 equal rows", flush=True)
    st, ost, ext, icp, on = _select(fm)
    _check(st, ost, ext, icp, on)


@pytest.mark.parametrize("m,seed", [(3, 30), (5, 31), (8, 32)])
def test_random_shells(m, seed):
    rng = np.random.default_rng(seed)
    fm = _sphere(rng, m, 2 * 500)
    print(f"normalization random m={m}", flush=True)
    st, ost, ext, icp, on = _select(fm, seed=seed)
    _check(st, ost, ext, icp, on)


def test_full_size_selection_input():
    """The standalone selection input at full size: 2N = 65536 rows s*u on the
    positive orthant, M = 3, K = 32640."""
    rng = np.random.default_rng(2024)
    fm = _sphere(rng, 3, 65536)
    print("normalization full-size selection input", flush=True)
    st, ost, ext, icp, on = _select(fm, seed=1)
    _check(st, ost, ext, icp, on)
