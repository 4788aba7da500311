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
