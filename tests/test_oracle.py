"""Pins the CPU oracle with checks that do not restate its formulas:
published known-answer vectors, closed forms, independent brute force and
structural invariants.  CPU only."""
import itertools
import math

import numpy as np
import pytest

import oracle as O


# ----------------------------------------------------------------- draws ----
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32-10
    assert list(O.philox4x32_10([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E,
                                                                         0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                [0xA4093822, 0x299F31D0])) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                               0x24126EA1]


def test_splitmix_known_answers():
    # splitmix64 seeded with 0: outputs are mix(k * golden), k = 1, 2, 3 (Vigna)
    g = 0x9E3779B97F4A7C15
    assert O.splitmix64_mix(g) == 0xE220A8397B1DCDAF
    assert O.splitmix64_mix((2 * g) % 2**64) == 0x6E789E6AA1B965F4
    assert O.splitmix64_mix((3 * g) % 2**64) == 0x06C45D188009454F
    # the reference's niching draw is the low word of mix(counter ^ seed)
    assert O.splitmix32(g, 0) == 0x7B1DCDAF
    assert O.splitmix32(5, 5) == O.splitmix64_mix(0) & 0xFFFFFFFF


def test_u01_range():
    assert O.u01(0) == 0.0
    assert O.u01(0xFFFFFFFF) == pytest.approx(1.0 - 2.0**-24, abs=0)
    assert O.u01(0xFFFFFFFF) < 1.0


# ------------------------------------------------------- elementary math ----
def _ulp32(x):
    x = np.float32(x)
    return float(np.spacing(np.abs(x)))


def test_sbx_pm_draws_uniform_and_independent():
    """The SBX draws and both children's mutation gates share one Philox block
    per element: every value must still be uniform on [0, 1) and the values
    of one element pairwise uncorrelated (a layout that reused bits would
    correlate two of them), the SBX direction a fair coin, and the deltas of
    rows 2p and 2p + 1 distinct streams."""
    pairs, d = 2000, 257
    ub, up, sg, uc = O.draw_sbx(12345, 7, pairs, d)
    ug, ud = O.draw_pm(12345, 7, 2 * pairs, d)
    ga = ug.reshape(pairs, 2, d)[:, 0].ravel()
    gb = ug.reshape(pairs, 2, d)[:, 1].ravel()
    da = ud.reshape(pairs, 2, d)[:, 0].ravel()
    db = ud.reshape(pairs, 2, d)[:, 1].ravel()
    vals = {"ub": ub, "up": up, "uc": uc, "ga": ga, "gb": gb, "da": da, "db": db}
    n = ub.size
    for name, v in vals.items():
        assert v.min() >= 0.0 and v.max() < 1.0, name
        assert abs(v.mean() - 0.5) < 4 * math.sqrt(1 / 12 / n), name
        hist, _ = np.histogram(v, bins=16, range=(0.0, 1.0))
        chi2 = ((hist - n / 16) ** 2 / (n / 16)).sum()
        assert chi2 < 50, (name, chi2)        # 15 dof: p < 1e-5 beyond 50
    names = list(vals)
    for a in range(len(names)):
        for b in range(a + 1, len(names)):
            r = np.corrcoef(vals[names[a]], vals[names[b]])[0, 1]
            assert abs(r) < 5 / math.sqrt(n), (names[a], names[b], r)
    assert abs(sg.mean() - 0.5) < 4 * math.sqrt(0.25 / n)
    # a small gate rate opens as often as it should
    rate = 1.0 / 64
    assert abs((gb < rate).mean() - rate) < 5 * math.sqrt(rate / n)
    assert abs((ga < rate).mean() - rate) < 5 * math.sqrt(rate / n)


def test_powf_close_to_fp64():
    rng = np.random.default_rng(0)
    # the SBX / PM call sites: base in [0, 2] or up to 2^24 with a root exponent,
    # base in [0, 2] with exponent eta + 1
    xs = np.concatenate([rng.random(2000) * 2.0, rng.random(500) * 2e6, rng.random(500) * 2.0,
                         [1e-30, 0.5, 1.5, 2.0]])
    ys = np.concatenate([1.0 / (1.0 + rng.random(2000) * 60), 1.0 / (1.0 + rng.random(500) * 60),
                         rng.random(500) * 40 + 1, [21.0, 1 / 31.0, 31.0, 0.5]])
    for x, y in zip(xs.astype(np.float32), ys.astype(np.float32)):
        exact = math.pow(float(x), float(y))
        got = O.powf(float(x), float(y))
        assert abs(got - exact) <= _ulp32(exact), (x, y, got, exact)
    assert O.powf(0.0, 0.5) == 0.0 and O.powf(1.0, 7.0) == 1.0


def test_powf_root_within_two_ulp_of_fp64():
    """The float32 root used by SBX and the mutation root, against the
    double-precision power: every SBX base 2u / (2 - 2u)^-1 for u on the
    2^-24 draw grid, and a dense sweep of x in [2^-30, 2] (the mutation root's
    vm), for several distribution indices."""
    def ulps(got, x, y):
        exact = x.astype(np.float64) ** np.float64(y)
        ref = exact.astype(np.float32)
        spacing = (np.nextafter(ref, np.float32(np.inf)) - ref).astype(np.float64)
        return np.abs(got.astype(np.float64) - exact) / spacing

    u = np.arange(1 << 24, dtype=np.float64) * 2.0 ** -24
    base = np.where(u < 0.5, 2.0 * u, 1.0 / (2.0 - 2.0 * u)).astype(np.float32)
    base = base[base > 0]
    sweep = np.arange(np.float32(2.0 ** -30).view(np.int32), np.float32(2.0).view(np.int32), 13,
                      dtype=np.int32).view(np.float32)
    worst = 0.0
    for eta in (0.5, 1.0, 20.0, 30.0):
        y = np.float32(1.0) / np.float32(1.0 + eta)
        for x in (base, sweep):
            worst = max(worst, float(ulps(O.powf_root(x, y), x, y).max()))
    assert worst <= 2.0, worst
    assert O.powf_root(0.0, 0.5) == 0.0 and O.powf_root(1.0, 0.25) == 1.0
    assert O.powf_root(2.0 ** -140, 0.5) == pytest.approx(2.0 ** -70, rel=2 ** -22)   # subnormal input
    assert np.isnan(O.powf_root(-1.0, 0.5)) and O.powf_root(float("inf"), 0.5) == float("inf")


def test_cosf_det_is_rounded_fp64_cos():
    """cos of a float32 through the fixed double sequence rounds the double
    cosine: equal to float32(cos(x)) (at most 1 ulp where the double result
    sits at a float32 rounding tie), over the DTLZ distance-term range
    |20 pi v| <= 10 pi, the float32 neighbours of k pi / 2 and a wider band."""
    rng = np.random.default_rng(7)
    xs = np.concatenate([
        rng.uniform(-10 * math.pi, 10 * math.pi, 200000),
        rng.uniform(-400.0, 400.0, 20000),
        [k * math.pi / 2 for k in range(-64, 65)],
        [0.0, -0.0, 1e-30, 2.0 ** -24, 1.0, 31.41592653589793],
    ]).astype(np.float32)
    got = O.cosf_det(xs).astype(np.float64)
    ref = np.cos(xs.astype(np.float64))
    refs = ref.astype(np.float32)
    spacing = np.abs(np.nextafter(refs, np.float32(np.inf)) - refs).astype(np.float64)
    spacing = np.maximum(spacing, 2.0 ** -149)
    err = np.abs(got - ref) / spacing
    assert err.max() <= 1.0, float(err.max())
    assert np.count_nonzero(got != refs.astype(np.float64)) <= 3
    assert O.cosf_det(0.0) == 1.0 and O.cosf_det(np.float32(math.pi)) == np.float32(-1.0)


def _dtlz1_fp64(x, m):
    x = x.astype(np.float64)
    xm, xk = x[:, : m - 1], x[:, m - 1:]
    g = 100.0 * (xk.shape[1] + ((xk - 0.5) ** 2 - np.cos(20 * math.pi * (xk - 0.5))).sum(1))
    f = np.empty((m, x.shape[0]))
    for i in range(m):
        v = 0.5 * (1 + g) * np.prod(xm[:, : m - 1 - i], axis=1)
        if i > 0:
            v = v * (1 - xm[:, m - 1 - i])
        f[i] = v
    return f, g


@pytest.mark.parametrize("n,d,m", [(300, 12, 3), (200, 140, 5), (64, 3, 3)])
def test_dtlz1_against_fp64_definition(n, d, m):
    """DTLZ1 (Deb, Thiele, Laumanns, Zitzler 2002) evaluated in float32 in the
    library's summation order, against the definition in double."""
    x = np.random.default_rng(n + d).random((n, d), dtype=np.float32)
    f, cv = O.eval_dtlz1(x, m)
    ref, _ = _dtlz1_fp64(x, m)
    np.testing.assert_allclose(f, ref, rtol=2e-5, atol=1e-5)
    assert not cv.any()


def test_dtlz1_front_and_c1_constraint():
    """On the Pareto front (tail variables 0.5) the objectives sum to 0.5 and
    C1-DTLZ1 is feasible there; at the corner x1 = x2 = 1 (f = (0.5(1+g), 0, 0))
    its violation is g, at x1 = 0 (f = (0, 0, 0.5(1+g))) it is
    max(0.5(1+g)/0.6 - 1, 0) (Jain and Deb 2014: the feasible band between the
    front and the plane through 0.5 on the first M-1 axes and 0.6 on the last)."""
    rng = np.random.default_rng(3)
    n, d, m = 256, 10, 3
    x = rng.random((n, d), dtype=np.float32)
    x[:, m - 1:] = 0.5
    f, cv = O.eval_dtlz1(x, m, constrained=True)
    np.testing.assert_allclose(f.sum(0), 0.5, rtol=1e-6)
    assert not cv.any()
    x = rng.random((n, d), dtype=np.float32) * np.float32(0.02) + np.float32(0.49)
    x[: n // 2, :2] = 1.0
    x[n // 2:, 0] = 0.0
    f, cv = O.eval_dtlz1(x, m, constrained=True)
    _, g = _dtlz1_fp64(x, m)
    np.testing.assert_allclose(cv[: n // 2], g[: n // 2], rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(cv[n // 2:], np.maximum(0.5 * (1 + g[n // 2:]) / 0.6 - 1, 0), rtol=1e-4,
                               atol=1e-4)
    assert (cv[: n // 2] > 0).all()


def test_acosf_close_to_fp64():
    for x in np.linspace(0, 1, 5001, dtype=np.float32):
        exact = math.acos(float(x))
        assert abs(O.acosf(float(x)) - exact) <= _ulp32(exact) + 1e-12, x
    assert O.acosf(1.0) == 0.0


# ------------------------------------------------------ Das-Dennis sets ----
@pytest.mark.parametrize("n,m,k", [(92, 3, 91), (32768, 3, 32640), (16384, 5, 14950), (210, 5, 210),
                                   (120, 8, 120), (100, 8, 72), (91, 3, 91), (90, 3, 78)])
def test_das_dennis_counts(n, m, k):
    assert O.das_dennis(n, m).shape == (m, k)


@pytest.mark.parametrize("h,m", [(12, 3), (6, 5), (3, 8), (4, 4)])
def test_das_dennis_order_matches_combinations(h, m):
    # prev_permutation over a mask whose first m-1 bits are set visits the
    # bit positions as (m-1)-combinations in lexicographic order
    _, raw = O.das_dennis(0, m, partitions=h, raw=True)
    expect = []
    for combo in itertools.combinations(range(h + m - 1), m - 1):
        pos = (-1,) + combo + (h + m - 1,)
        expect.append([(pos[i + 1] - pos[i] - 1) / h for i in range(m)])
    expect = np.array(expect, np.float32).T
    np.testing.assert_array_equal(raw, expect)
    assert np.allclose(raw.sum(axis=0), 1.0, atol=1e-6)


def test_das_dennis_unit_and_two_layer():
    w, raw = O.das_dennis(100, 8, raw=True)
    assert np.allclose(np.linalg.norm(w.astype(np.float64), axis=0), 1.0, atol=1e-6)
    # second layer: 0.5 * (p + 1/M), inside the simplex
    assert np.allclose(raw[:, 36:].sum(axis=0), 1.0, atol=1e-6)
    assert raw[:, 36:].min() >= 0.5 / 8 - 1e-7
    w2, raw2 = O.das_dennis(16384, 5, kind=2, raw=True)
    assert raw2.min() >= 1e-6


# ------------------------------------------------------------ dominance ----
def brute_fronts(f, cv, stop_at):
    """Independent definition-level peeling (plain Python)."""
    m, n = f.shape

    def dom(q, p):
        if q == p:
            return False
        fq, fp = cv[q] == 0.0, cv[p] == 0.0
        if fq and not fp:
            return True
        if not fq and not fp:
            return bool(cv[q] < cv[p])
        if not fq and fp:
            return False
        return all(f[k, q] <= f[k, p] for k in range(m)) and any(f[k, q] < f[k, p] for k in range(m))

    fronts = [-1] * n
    remaining = set(range(n))
    front, assigned, last = 0, 0, 0
    while remaining and assigned < stop_at:
        cur = [p for p in remaining if not any(dom(q, p) for q in remaining)]
        for p in cur:
            fronts[p] = front
        remaining -= set(cur)
        assigned += len(cur)
        last = front
        front += 1
    return np.array(fronts, np.int32), last


@pytest.mark.parametrize("seed", range(8))
def test_nds_brute_force(seed):
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(1, 40)), int(rng.integers(1, 5))
    f = (rng.integers(0, 4, (m, n)) / 4.0).astype(np.float32)
    cv = np.where(rng.random(n) < 0.3, rng.integers(0, 3, n) * 0.5, 0.0).astype(np.float32)
    if seed % 3 == 0:
        f[:, rng.random(n) < 0.1] = np.nan
    for stop in (n, max(1, n // 2), 1):
        fr, mf = O.nds(f, cv, stop)
        bfr, bmf = brute_fronts(f, cv, stop)
        np.testing.assert_array_equal(fr, bfr)
        assert mf == bmf


def test_nds_invariants_large():
    rng = np.random.default_rng(3)
    n, m = 1500, 3
    f = rng.random((m, n), dtype=np.float32)
    cv = np.zeros(n, np.float32)
    fr, mf = O.nds(f, cv, n)
    assert fr.min() == 0 and fr.max() == mf
    for k in range(1, mf + 1):
        prev = np.where(fr == k - 1)[0]
        for p in np.where(fr == k)[0][:20]:
            # some member of front k-1 dominates p, none of front k does
            le = (f[:, prev] <= f[:, [p]]).all(axis=0) & (f[:, prev] < f[:, [p]]).any(axis=0)
            assert le.any()
            same = np.where(fr == k)[0]
            le2 = (f[:, same] <= f[:, [p]]).all(axis=0) & (f[:, same] < f[:, [p]]).any(axis=0)
            assert not le2.any()


# ----------------------------------------------------------- hyperplane ----
def test_hyperplane_closed_forms():
    ill, a = O.hyperplane_fp64(np.eye(3))
    assert ill == 0 and np.allclose(a, 1.0, atol=1e-15)
    ill, a = O.hyperplane_fp64(np.diag([2.0, 4.0, 8.0]))
    assert ill == 0 and np.allclose(a, [0.5, 0.25, 0.125], rtol=1e-14)
    # duplicate extreme points -> singular -> fallback
    ill, _ = O.hyperplane_fp64(np.array([[1.0, 0, 0], [1.0, 0, 0], [0, 0, 1.0]]))
    assert ill == 1
    # condition number above 1e6 -> fallback
    ill, _ = O.hyperplane_fp64(np.diag([1.0, 1.0, 1e-7]))
    assert ill == 1
    ill, _ = O.hyperplane_fp64(np.array([[np.inf, 0, 0], [0, 1, 0], [0, 0, 1.0]]))
    assert ill == 1


@pytest.mark.parametrize("m,seed", [(2, 0), (3, 1), (5, 2), (8, 3), (10, 4)])
def test_hyperplane_matches_lapack(m, seed):
    rng = np.random.default_rng(seed)
    E = np.eye(m) * 2 + rng.random((m, m)) * 0.3
    ill, a = O.hyperplane_fp64(E)
    assert ill == 0
    np.testing.assert_allclose(a, np.linalg.solve(E, np.ones(m)), rtol=1e-12)


def test_hyperplane_equal_rows_always_fall_back():
    """Two equal rows of E (one individual extreme on two axes) always end in
    the row-max fallback: the device takes it without sweeping, which is
    exact only if the full solve never accepts such a matrix."""
    rng = np.random.default_rng(11)
    for m in range(2, 11):
        for trial in range(40):
            scale = 10.0 ** rng.uniform(-6, 6)
            E = (np.abs(rng.normal(0, 0.3, (m, m))) + np.eye(m)) * scale
            i, j = rng.choice(m, 2, replace=False)
            E[j] = E[i]
            ill, _ = O.hyperplane_fp64(E)
            assert ill == 1, (m, trial)


def test_hyperplane_truncated_pinv():
    # sigma_min in (1e-10, 1e-5] with kappa <= 1e6: that singular value is dropped
    U, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((3, 3)))
    V, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))
    s = np.array([5.0, 1.0, 8e-6])
    E = U @ np.diag(s) @ V.T
    ill, a = O.hyperplane_fp64(E)
    assert ill == 0
    u, sv, vt = np.linalg.svd(E)
    inv = np.where(sv > 1e-5, 1.0 / sv, 0.0)
    np.testing.assert_allclose(a, vt.T @ (inv * (u.T @ np.ones(3))), rtol=1e-9, atol=1e-12)


def test_normalize_fallback_and_nan():
    # all points on one axis: duplicate extremes -> row-max fallback
    f = np.array([[1.0, 2.0, 4.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
    fn, a, ext, a64, ill = O.normalize(f, np.zeros(3, np.float32))
    assert ill == 1
    np.testing.assert_array_equal(a, np.array([0.25, 1.0, 1.0], np.float32))
    # identity extremes: intercepts 1
    f = np.array([[3.0, 1.0, 1.0, 2.0], [1.0, 3.0, 1.0, 2.0], [1.0, 1.0, 3.0, 2.0]], np.float32)
    fn, a, ext, a64, ill = O.normalize(f, np.ones(3, np.float32))
    assert ill == 0 and list(ext) == [0, 1, 2]
    np.testing.assert_allclose(a64, [0.5, 0.5, 0.5], rtol=1e-14)
    # a NaN objective is never an extreme point and never part of the row max
    f = np.array([[np.nan, 1.0, 0.0], [0.0, 0.0, 1.0]], np.float32)
    fn, a, ext, a64, ill = O.normalize(f, np.zeros(2, np.float32))
    assert 0 not in list(ext)


# ---------------------------------------------------------- association ----
def test_association_matches_fp64_argmin():
    rng = np.random.default_rng(0)
    w = O.das_dennis(120, 3)
    x = rng.random((3, 500), dtype=np.float32)
    dist, ref = O.associate(x, w)
    X, W = x.astype(np.float64), w.astype(np.float64)
    dots = X.T @ W
    d2 = (X * X).sum(0)[:, None] - 2 * dots ** 2 + dots ** 2 * (W * W).sum(0)[None, :]
    best = d2.argmin(1)
    gap = np.sort(d2, 1)[:, 1] - np.sort(d2, 1)[:, 0]
    clear = gap > 1e-5
    np.testing.assert_array_equal(ref[clear], best[clear])
    np.testing.assert_allclose(dist, np.sqrt(np.maximum(d2.min(1), 0)), atol=2e-4)


def test_association_ties_take_smallest_index():
    w = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, 0.0]], np.float32)  # directions 0 and 2 identical
    x = np.array([[2.0, 0.0], [0.0, 3.0]], np.float32)
    dist, ref = O.associate(x, w)
    assert list(ref) == [0, 1]
    x = np.array([[np.nan], [1.0]], np.float32)
    dist, ref = O.associate(x, w)
    assert ref[0] == -1 and np.isinf(dist[0])


# -------------------------------------------------------------- niching ----
def test_niching_hand_case():
    # 2 prior members on ref 0; last front: refs [0, 1, 1, 2] with distances
    ref = np.array([0, 0, 0, 1, 1, 2], np.int32)
    dist = np.array([0.1, 0.2, 0.05, 0.3, 0.1, 0.9], np.float32)
    ua = np.zeros(4, np.uint32)
    ub = np.zeros(4, np.uint32)
    # empty niches 1 and 2 take their closest member first (key 0): local 4 (d=.1) and 5;
    # then round rho+rank: ref1 member 3 (round 1), ref0 member 2 (round 2)
    out = O.niching(ref, dist, 2, 4, 3, 2, ua, ub)
    assert sorted(out.tolist()) == [4, 5]
    out = O.niching(ref, dist, 2, 4, 3, 3, ua, ub)
    assert sorted(out.tolist()) == [3, 4, 5]
    out = O.niching(ref, dist, 2, 4, 3, 4, ua, ub)
    assert sorted(out.tolist()) == [2, 3, 4, 5]


def test_niching_invariants():
    rng = np.random.default_rng(1)
    k, n_prior, n_last, n_rem = 30, 40, 60, 25
    ref = rng.integers(0, k, n_prior + n_last).astype(np.int32)
    dist = rng.random(n_prior + n_last, dtype=np.float32)
    ua = rng.integers(0, 2**32, n_last, dtype=np.uint64).astype(np.uint32)
    ub = rng.integers(0, 2**32, n_last, dtype=np.uint64).astype(np.uint32)
    out = O.niching(ref, dist, n_prior, n_last, k, n_rem, ua, ub)
    assert len(set(out.tolist())) == n_rem and out.min() >= n_prior
    rho = np.bincount(ref[:n_prior], minlength=k)
    # every empty niche with a last-front member contributes its closest member
    # as long as there is room (virtual round 0)
    empty = [j for j in range(k) if rho[j] == 0 and (ref[n_prior:] == j).any()]
    if len(empty) <= n_rem:
        for j in empty:
            members = np.where(ref[n_prior:] == j)[0] + n_prior
            closest = members[np.argmin(dist[members])]
            assert closest in out


# ---------------------------------------------------------- reproduction ----
def test_tournament_rules():
    rank = np.array([0, 1, 2, 0], np.int32)
    cv = np.array([0.0, 0.0, 0.5, 0.2], np.float32)
    w = O.tournament([0, 1, 2, 2, 0], [1, 0, 3, 0, 0], rank, cv)
    assert list(w) == [0, 0, 3, 0, 0]


def test_sbx_closed_forms():
    d = 3
    parents = np.array([[0.2, 0.4, 0.6], [0.8, 0.4, 0.1]], np.float32)
    bounds = np.tile(np.array([0.0, 1.0], np.float32), d)
    idx = np.array([0, 1], np.int32)
    half = np.full(d, 0.5, np.float32)
    # u = 1/2 -> beta = 1: children are the parents (order given by the sign)
    off = O.sbx(parents, idx, bounds, half, np.zeros(d, np.float32), np.ones(d, np.int32),
                np.ones(d, np.float32), 30.0, 1.0, 0.0, 2)
    np.testing.assert_allclose(off, parents, atol=1e-7)
    # probability gate closed -> copy
    off = O.sbx(parents, idx, bounds, np.full(d, 0.1, np.float32), np.full(d, 0.99, np.float32),
                np.zeros(d, np.int32), np.ones(d, np.float32), 30.0, 0.5, 0.0, 2)
    np.testing.assert_array_equal(off, parents)


def test_sbx_against_fp64_formula():
    rng = np.random.default_rng(2)
    n, d, eta = 40, 9, 15.0
    parents = rng.random((n, d), dtype=np.float32)
    bounds = np.tile(np.array([0.0, 1.0], np.float32), d)
    idx = rng.integers(0, n, n).astype(np.int32)
    ub, sg = rng.random(n // 2 * d, dtype=np.float32), rng.integers(0, 2, n // 2 * d).astype(np.int32)
    off = O.sbx(parents, idx, bounds, ub, np.zeros_like(ub), sg, np.ones_like(ub), eta, 1.0, 0.0, n)
    for pr in range(n // 2):
        for j in range(d):
            e = pr * d + j
            u = float(ub[e])
            beta = (2 * u) ** (1 / (1 + eta)) if u < 0.5 else (1 / (2 - 2 * u)) ** (1 / (1 + eta))
            beta *= 2 * sg[e] - 1
            x1, x2 = float(parents[idx[2 * pr], j]), float(parents[idx[2 * pr + 1], j])
            ca = min(max(0.5 * (x1 + x2) + 0.5 * beta * (x1 - x2), 0.0), 1.0)
            # float32 rounding of mid and delta: a few ulp of the operand scale
            scale = max(abs(x1), abs(x2), abs(ca))
            assert abs(off[2 * pr, j] - ca) <= 3 * _ulp32(scale)


def test_pm_identity_and_bounds():
    rng = np.random.default_rng(3)
    n, d = 20, 6
    x = rng.random((n, d), dtype=np.float32)
    bounds = np.tile(np.array([0.0, 1.0], np.float32), d)
    out = O.pm(x, bounds, np.ones(n * d, np.float32) * 0.9, rng.random(n * d, dtype=np.float32), 20.0, 1.0)
    np.testing.assert_array_equal(out, x)  # gate 0.9 >= 1/6 -> unchanged
    out = O.pm(x, bounds, np.zeros(n * d, np.float32), rng.random(n * d, dtype=np.float32), 20.0, 1.0)
    assert out.min() >= 0.0 and out.max() <= 1.0 and not np.array_equal(out, x)


def test_annealing():
    assert O.annealed(30.0, 10.0, 0, 100) == 30.0
    assert O.annealed(30.0, 10.0, 50, 100) == 20.0
    assert O.annealed(30.0, 10.0, 500, 100) == 10.0


# ------------------------------------------------------------ objectives ----
def test_quadratic_close_to_fp64():
    rng = np.random.default_rng(4)
    x = rng.random((50, 1024), dtype=np.float32)
    c = rng.random((3, 1024), dtype=np.float32)
    f = O.eval_quadratic(x, c)
    ref = ((x[None, :, :].astype(np.float64) - c[:, None, :]) ** 2).sum(-1)
    np.testing.assert_allclose(f, ref, rtol=2e-6)


# ----------------------------------------------------------------- RVEA ----
def test_rvea_gamma_closed_form():
    # orthonormal axes: nearest neighbour at 90 degrees
    g = O.rvea_gamma(np.eye(3, dtype=np.float32))
    np.testing.assert_allclose(g, np.pi / 2, rtol=1e-6)
    v = O.das_dennis(91, 3)
    g = O.rvea_gamma(v)
    V = v.astype(np.float64)
    cos = V.T @ V
    np.fill_diagonal(cos, -np.inf)
    np.testing.assert_allclose(g, np.arccos(np.clip(cos.max(1), 0, 1)), atol=2e-6)


def test_rvea_select_one_winner_per_direction():
    rng = np.random.default_rng(5)
    n, m = 60, 3
    v = O.das_dennis(n, m, kind=2)
    gam = O.rvea_gamma(v)
    fm = rng.random((m, 2 * n), dtype=np.float32) + 0.1
    cvm = np.zeros(2 * n, np.float32)
    sel, na, ref, th, apd = O.rvea_select(fm, cvm, v, gam, n, O.rvea_penalty(m, 5, 100, 2.0))
    assert na == len(sel) <= min(n, v.shape[1])
    assert len(set(ref[sel].tolist())) == len(sel)          # one per direction
    for r in set(ref.tolist()):
        members = np.where(ref == r)[0]
        assert apd[members].min() == apd[sel[ref[sel] == r]][0]  # the APD minimum wins


def test_rvea_adapt_identity_range():
    v0 = O.das_dennis(91, 3, kind=2)
    f = np.stack([np.linspace(0, 1, 50)] * 3).astype(np.float32)
    out = O.rvea_adapt(v0, f, 50)
    np.testing.assert_allclose(out, v0, atol=1e-7)


# --------------------------------------------- RVEA against its definition ----
def _rvea_fp64(fm, cvm, v, gam, active, penalty):
    """The published RVEA selection (Cheng et al., IEEE TEVC 2016, with the
    reference's conventions: translation by the minimum over all 2n columns,
    floor 1e-32, dead parent slots excluded, feasible rows by APD, one
    survivor per direction, truncation to n by APD) in float64, written from
    the definition: per direction, the member with the smallest APD survives."""
    m, nm = fm.shape
    n = nm // 2
    F = fm.astype(np.float64)
    zmin = np.min(F, axis=1, keepdims=True)
    T = np.maximum(F - zmin, 1e-32)
    V = v.astype(np.float64)
    nrm = np.sqrt((T * T).sum(0))
    cos = (V.T @ T) / np.sqrt((T * T).sum(0) + 1e-12)          # (k, 2n)
    ref = np.argmax(cos, axis=0)
    c = np.clip(cos[ref, np.arange(nm)], 0.0, 1.0)
    theta = np.arccos(c)
    ref = ref.astype(np.int64)
    ref[active:n] = -1                                          # dead parent slots
    apd = (1.0 + penalty * np.clip(theta, 0, np.pi) / np.maximum(gam.astype(np.float64)[np.maximum(ref, 0)],
                                                                     1e-6)) * nrm
    winners = []
    for j in range(v.shape[1]):
        members = np.where(ref == j)[0]
        if len(members):
            winners.append(members[np.argmin(apd[members])])
    winners = np.array(sorted(winners, key=lambda i: apd[i]))[:n]
    return np.sort(winners), ref, theta, c, apd


@pytest.mark.parametrize("n,m,active,seed", [(30, 2, 30, 0), (60, 3, 60, 1), (60, 3, 41, 2), (105, 5, 105, 3),
                                             (120, 8, 77, 4), (50, 3, 50, 5), (200, 4, 180, 6)])
def test_rvea_select_against_fp64_definition(n, m, active, seed):
    rng = np.random.default_rng(seed)
    v = O.das_dennis(n, m, kind=2)
    gam = O.rvea_gamma(v)
    fm = (rng.random((m, 2 * n)) * 3 + 0.25).astype(np.float32)
    fm[:, active:n] = 1e15
    cvm = np.zeros(2 * n, np.float32)
    cvm[active:n] = 1e15
    pen = O.rvea_penalty(m, 37, 100, 2.0)
    sel, na, ref, th, apd = O.rvea_select(fm, cvm, v, gam, active, pen)
    w64, ref64, th64, c64, apd64 = _rvea_fp64(fm, cvm, v, gam, active, pen)
    live = np.r_[0:active, n:2 * n]
    np.testing.assert_array_equal(ref[live], ref64[live])
    assert np.all(ref[active:n] == v.shape[1])
    # theta: the float32 cosine carries a few ulp of 1 (an M-term fma chain and
    # a reciprocal square root), which acos amplifies by 1 / sqrt(1 - c^2)
    ulp1 = 2.0 ** -24
    tol_th = 4 * ulp1 * np.pi + 8 * m * ulp1 / np.sqrt(np.maximum(1 - c64[live] ** 2, 1e-12))
    assert np.all(np.abs(th[live] - th64[live]) <= tol_th)
    # APD = (1 + P theta / gamma) ||f'||: relative error of the norm plus the
    # theta error scaled by P / gamma
    g = np.maximum(gam.astype(np.float64)[ref64[live]], 1e-6)
    nrm = apd64[live] / (1 + pen * np.clip(th64[live], 0, np.pi) / g)
    tol_apd = 8 * ulp1 * apd64[live] + pen / g * tol_th * nrm
    assert np.all(np.abs(apd[live] - apd64[live]) <= tol_apd)
    np.testing.assert_array_equal(sel, w64)
    assert na == len(w64)


def test_rvea_penalty_closed_forms():
    # at t = G the penalty is exactly M; at t = 0 it is 0; alpha = 1 is linear
    for m in (2, 3, 5, 10):
        assert O.rvea_penalty(m, 100, 100, 2.0) == float(m)
        assert O.rvea_penalty(m, 0, 100, 2.0) == 0.0
        assert O.rvea_penalty(m, 250, 100, 2.0) == float(m)          # progress clamps at 1
    assert O.rvea_penalty(3, 25, 100, 1.0) == np.float32(0.75)
    assert O.rvea_penalty(4, 50, 100, 2.0) == 1.0


def test_rvea_two_directions_known_angle():
    """Two directions 60 degrees apart, one individual 10 degrees from the
    first: it belongs to the first, theta = 10 degrees, and the APD is
    (1 + P theta / gamma) |f'| with gamma = 60 degrees."""
    phi, psi = math.radians(60), math.radians(10)
    v = np.array([[1.0, math.cos(phi)], [0.0, math.sin(phi)]], np.float32)
    gam = O.rvea_gamma(v)
    np.testing.assert_allclose(gam, [phi, phi], rtol=1e-6)
    # merged population of 2 (n = 1): the anchor at the origin after translation,
    # and the individual at radius 2 and angle psi from direction 0
    fm = np.array([[0.0, 2 * math.cos(psi)], [0.0, 2 * math.sin(psi)]], np.float32)
    pen = 2.0
    sel, na, ref, th, apd = O.rvea_select(fm, np.zeros(2, np.float32), v, gam, 1, pen)
    assert ref[1] == 0
    np.testing.assert_allclose(th[1], psi, rtol=1e-5)
    np.testing.assert_allclose(apd[1], (1 + pen * psi / phi) * 2.0, rtol=1e-5)
    assert apd[0] == 0.0                                       # zero norm after translation


def test_rvea_adapt_shifted_objectives_and_dead_rows():
    """V = normalize(V0 * (zmax - zmin)) over the live rows only: objectives
    shifted by +5 (so zmin != 0 matters) and dead rows at 1e15 past n_active."""
    rng = np.random.default_rng(7)
    m, n, active = 3, 80, 53
    v0 = O.das_dennis(91, m, kind=2)
    f = (rng.random((m, n)) * np.array([[1.0], [3.0], [0.5]]) + 5.0).astype(np.float32)
    f[:, active:] = 1e15
    out = O.rvea_adapt(v0, f, active)
    rng_ = f[:, :active].astype(np.float64).max(1) - f[:, :active].astype(np.float64).min(1)
    s = v0.astype(np.float64) * rng_[:, None]
    ref = s / np.linalg.norm(s, axis=0, keepdims=True)
    np.testing.assert_allclose(out, ref, rtol=0, atol=8 * 2.0 ** -24)
    # a dropped "- zmin" or a range over all n rows would be far off
    wrong = v0.astype(np.float64) * f[:, :active].max(1)[:, None]
    assert np.abs(out - wrong / np.linalg.norm(wrong, axis=0)).max() > 1e-2


# ---------------------------------------------- PM against Deb's formula ----
def test_pm_against_fp64_deb_formula():
    """Polynomial mutation (Deb & Goyal 1996, in the bounded form the reference
    uses): with x' = (x - lo) / (hi - lo), for u <= 1/2
    delta = (2u + (1 - 2u)(1 - x')^(eta+1))^(1/(eta+1)) - 1, else
    delta = 1 - (2(1 - u) + 2(u - 1/2) x'^(eta+1))^(1/(eta+1)); the child is
    clamp(x + delta (hi - lo)).  Both branches, values near both bounds."""
    rng = np.random.default_rng(8)
    n, d, eta = 64, 8, 20.0
    lo = np.array([-1.0, 0.0, 2.0, -5.0, 0.0, 0.0, 10.0, -0.5], np.float32)
    hi = np.array([1.0, 1.0, 3.0, 5.0, 0.001, 100.0, 10.5, 0.5], np.float32)
    bounds = np.stack([lo, hi], 1).reshape(-1).astype(np.float32)
    t = rng.random((n, d))
    t[:8] = 1e-4          # near the lower bound
    t[8:16] = 1 - 1e-4    # near the upper bound
    x = (lo + t * (hi - lo)).astype(np.float32)
    ud = rng.random(n * d).astype(np.float32)
    ud[::3] = np.float32(0.5)                               # the branch point itself
    out = O.pm(x, bounds, np.zeros(n * d, np.float32), ud, eta, 1.0)
    for r in range(n):
        for j in range(d):
            u = float(ud[r * d + j])
            xv, a, b = float(x[r, j]), float(lo[j]), float(hi[j])
            xn = (xv - a) / (b - a)
            e1 = eta + 1.0
            if u <= 0.5:
                delta = (2 * u + (1 - 2 * u) * (1 - xn) ** e1) ** (1 / e1) - 1
            else:
                delta = 1 - (2 * (1 - u) + 2 * (u - 0.5) * xn ** e1) ** (1 / e1)
            ref = min(max(xv + delta * (b - a), a), b)
            # float32 evaluation: the normalized value, two powf (within 1 ulp
            # each) and the fused update, a few ulp of the range
            assert abs(out[r, j] - ref) <= 16 * 2.0 ** -24 * (b - a) + _ulp32(ref), (r, j, u, out[r, j], ref)


# ------------------------------------------ niching against its definition ----
def _niching_definition(ref, dist, n_prior, n_last, k, n_rem, ua, ub):
    """NSGA-III niching with the reference's random tie-breaks, written from
    the definition (niching.cu:114-178) rather than as three sorts: list the
    last front by (direction, distance, index) and number that order p1;
    inside each direction, a direction no prior-front member uses (rho = 0)
    takes its closest member first (distance, then index), and every other
    member follows by the draw ua[p1] | 1 (then p1); a member's round is rho
    plus its place in that order.  Number the members by (direction, place)
    as p2; the survivors are the first n_rem by (round, ub[p2], p2)."""
    rho = np.zeros(k, np.int64)
    for r in ref[:n_prior]:
        if 0 <= r < k:
            rho[r] += 1
    members = list(range(n_prior, n_prior + n_last))
    dbits = lambda i: int(np.float32(dist[i]).view(np.uint32))
    p1_order = sorted(members, key=lambda i: (int(np.uint32(np.int64(ref[i]) & 0xFFFFFFFF)), dbits(i), i))
    p1 = {i: t for t, i in enumerate(p1_order)}
    by_dir = {}
    for i in p1_order:
        by_dir.setdefault(int(ref[i]), []).append(i)
    placed = []
    for r in sorted(by_dir, key=lambda r: r & 0xFFFFFFFF):
        lst = by_dir[r]
        rh = rho[r] if 0 <= r < k else 0
        first = [lst[0]] if rh == 0 else []          # lst is already by (distance, index)
        rest = sorted([i for i in lst if i not in first], key=lambda i: (int(ua[p1[i]]) | 1, p1[i]))
        for place, i in enumerate(first + rest):
            placed.append((r, place, i, rh + place))
    p2 = {i: t for t, (r, place, i, rnd) in enumerate(placed)}
    final = sorted(placed, key=lambda e: (e[3], int(ub[p2[e[2]]]), p2[e[2]]))
    return np.array([e[2] for e in final[:n_rem]], np.int32)


@pytest.mark.parametrize("case", range(24))
def test_niching_against_definition_with_draws(case):
    rng = np.random.default_rng(100 + case)
    k = int(rng.integers(2, 12))
    n_prior = int(rng.integers(0, 15))
    n_last = int(rng.integers(1, 25))
    n_rem = int(rng.integers(1, n_last + 1))
    ref = rng.integers(0, k, n_prior + n_last).astype(np.int32)
    # coarse distances make exact ties on purpose (the index breaks them)
    dist = (rng.integers(0, 6, n_prior + n_last) / 4.0).astype(np.float32)
    ua = rng.integers(0, 2 ** 32, n_last, dtype=np.uint64).astype(np.uint32)
    ub = rng.integers(0, 2 ** 32, n_last, dtype=np.uint64).astype(np.uint32)
    if case % 4 == 0:
        ub[:] = rng.integers(0, 3, n_last).astype(np.uint32)   # ties in the last key too
    got = O.niching(ref, dist, n_prior, n_last, k, n_rem, ua, ub)
    want = _niching_definition(ref, dist, n_prior, n_last, k, n_rem, ua, ub)
    np.testing.assert_array_equal(got, want)


def test_rvea_partition_ties_take_smallest_index():
    """Equal cosines to two directions (the bisector of the axes; a duplicated
    direction) go to the smaller direction index."""
    v = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, 0.0]], np.float32)       # directions 0 and 2 identical
    gam = np.full(3, 0.5, np.float32)
    fm = np.array([[0.0, 3.0, 1.0, 2.0], [0.0, 3.0, 0.0, 0.0]], np.float32)
    sel, na, ref, th, apd = O.rvea_select(fm, np.zeros(4, np.float32), v, gam, 2, 1.0)
    assert ref[1] == 0            # (3, 3) after translation: cos 1/sqrt2 to directions 0 and 1
    assert ref[2] == 0 and ref[3] == 0   # on the axis shared by directions 0 and 2
