"""GPU parity of the NSGA-III hot path: libmoea_hip.so (through the C ABI)
against the CPU oracle on the same inputs.  Integer/index results bit-exact;
floats produced by the same operation sequence bit-exact."""
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


def tie_objectives(rng, m, n, levels=7):
    """Objectives on a coarse grid: many exact ties and duplicate rows."""
    return rng.integers(0, levels, size=(m, n)).astype(np.float32) / np.float32(levels)


def sphere_objectives(rng, m, n):
    """Rows s*u, u uniform on the positive unit-sphere orthant, s in [1, 2]."""
    g = np.abs(rng.standard_normal((n, m)))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    s = rng.uniform(1.0, 2.0, size=(n, 1))
    return np.ascontiguousarray((s * u).T.astype(np.float32))


CASES = [(1, 1, 0), (2, 5, 1), (63, 3, 2), (64, 3, 3), (65, 2, 4), (184, 3, 5), (300, 5, 6),
         (1000, 8, 7), (1500, 10, 8), (4096, 3, 9)]


@pytest.mark.parametrize("n,m,seed", CASES)
def test_nds_random(n, m, seed):
    rng = np.random.default_rng(seed)
    f = tie_objectives(rng, m, n) if seed % 2 else rng.random((m, n), dtype=np.float32)
    cv = np.zeros(n, np.float32)
    for stop in (n, max(1, n // 2)):
        print(f"nds case n={n} m={m} seed={seed} stop={stop}", flush=True)
        fr_o, mf_o = O.nds(f, cv, stop)
        fr_g, mf_g = K.nds_rank(_f(f), _f(cv), stop)
        assert mf_g == mf_o
        np.testing.assert_array_equal(fr_g.cpu().numpy(), fr_o)


@pytest.mark.parametrize("n,m,seed", [(200, 3, 1), (777, 4, 2), (130, 2, 3)])
def test_nds_constraints_and_nan(n, m, seed):
    rng = np.random.default_rng(seed)
    f = tie_objectives(rng, m, n)
    cv = np.where(rng.random(n) < 0.3, rng.integers(0, 4, n).astype(np.float32) * 0.5, 0.0)
    cv = cv.astype(np.float32)
    f[:, rng.random(n) < 0.05] = np.nan
    f[0, rng.random(n) < 0.05] = np.inf
    cv[rng.random(n) < 0.03] = np.nan
    print(f"nds constrained case n={n} m={m} seed={seed}", flush=True)
    fr_o, mf_o = O.nds(f, cv, n)
    fr_g, mf_g = K.nds_rank(_f(f), _f(cv), n)
    assert mf_g == mf_o
    np.testing.assert_array_equal(fr_g.cpu().numpy(), fr_o)


@pytest.mark.parametrize("n,m,levels0,seed", [(1000, 3, 2, 11), (3000, 2, 3, 12), (777, 5, 1, 13),
                                               (2100, 3, 40, 14), (640, 1, 3, 15)])
def test_nds_f0_tie_groups(n, m, levels0, seed):
    """f0 on very few levels: tie groups of equal f0 run across many 64-row
    blocks, where a later row in sort order can dominate an earlier one."""
    rng = np.random.default_rng(seed)
    f = rng.random((m, n), dtype=np.float32)
    f[0] = rng.integers(0, levels0, n).astype(np.float32)
    if m > 1:
        f[1, rng.random(n) < 0.2] = 0.5                    # ties in other objectives too
    f[0, (f[0] == 0) & (rng.random(n) < 0.5)] = -0.0    # -0 and +0 compare equal
    cv = np.zeros(n, np.float32)
    for stop in (n, n // 3):
        print(f"nds tie case n={n} m={m} levels0={levels0} seed={seed} stop={stop}", flush=True)
        fr_o, mf_o = O.nds(f, cv, stop)
        fr_g, mf_g = K.nds_rank(_f(f), _f(cv), stop)
        assert mf_g == mf_o
        np.testing.assert_array_equal(fr_g.cpu().numpy(), fr_o)


def test_nds_duplicates_and_segments():
    """Duplicate rows, feasible rows with NaN objectives, infeasible rows with
    equal / NaN cv and +-inf objectives, interleaved across blocks."""
    rng = np.random.default_rng(21)
    n, m = 2500, 3
    base = rng.integers(0, 5, size=(m, 60)).astype(np.float32) / np.float32(5)
    f = base[:, rng.integers(0, 60, n)]
    f = np.ascontiguousarray(f)
    f[rng.integers(0, m, 40), rng.integers(0, n, 40)] = np.nan
    f[1, rng.random(n) < 0.02] = -np.inf
    f[2, rng.random(n) < 0.02] = np.inf
    cv = np.where(rng.random(n) < 0.4, rng.integers(1, 4, n).astype(np.float32), 0.0).astype(np.float32)
    cv[rng.random(n) < 0.02] = np.nan
    cv[rng.random(n) < 0.05] = -0.0                     # feasible (== 0)
    print("nds duplicates/segments case n=2500", flush=True)
    for stop in (n, 700):
        fr_o, mf_o = O.nds(f, cv, stop)
        fr_g, mf_g = K.nds_rank(_f(f), _f(cv), stop)
        assert mf_g == mf_o
        np.testing.assert_array_equal(fr_g.cpu().numpy(), fr_o)


@pytest.mark.parametrize("m", [3, 5, 8])
def test_nds_wide_sorts_with_ties_constraints_and_nan(m):
    """Sorts of 64 row blocks and more (the dominance kernel with its general
    tiles out of line up to six objectives, row ends from four): f0 on few
    levels so tie groups cross blocks, infeasible rows with equal and NaN cv,
    NaN and infinite objectives."""
    rng = np.random.default_rng(40 + m)
    n = 6000
    f = rng.random((m, n), dtype=np.float32)
    f[0] = rng.integers(0, 30, n).astype(np.float32) / np.float32(30)
    f[1, rng.random(n) < 0.2] = 0.5
    f[:, rng.random(n) < 0.01] = np.nan
    f[m - 1, rng.random(n) < 0.01] = np.inf
    cv = np.where(rng.random(n) < 0.2, rng.integers(1, 5, n).astype(np.float32) * 0.25, 0.0).astype(np.float32)
    cv[rng.random(n) < 0.01] = np.nan
    for stop in (n, n // 3):
        print(f"nds wide case n={n} m={m} stop={stop}", flush=True)
        fr_o, mf_o = O.nds(f, cv, stop)
        fr_g, mf_g = K.nds_rank(_f(f), _f(cv), stop)
        assert mf_g == mf_o
        np.testing.assert_array_equal(fr_g.cpu().numpy(), fr_o)


def test_nds_larger_random():
    rng = np.random.default_rng(31)
    n, m = 20000, 3
    f = rng.random((m, n), dtype=np.float32)
    cv = np.zeros(n, np.float32)
    print("nds random case n=20000", flush=True)
    fr_o, mf_o = O.nds(f, cv, n // 2)
    fr_g, mf_g = K.nds_rank(_f(f), _f(cv), n // 2)
    assert mf_g == mf_o
    np.testing.assert_array_equal(fr_g.cpu().numpy(), fr_o)


def test_nds_selection_bench_shape():
    """Standalone selection bench input (2N x 3 rows s*u) at N = 4096."""
    rng = np.random.default_rng(123)
    n = 8192
    f = sphere_objectives(rng, 3, n)
    cv = np.zeros(n, np.float32)
    print("nds sphere case n=8192", flush=True)
    fr_o, mf_o = O.nds(f, cv, n // 2)
    fr_g, mf_g = K.nds_rank(_f(f), _f(cv), n // 2)
    assert mf_g == mf_o
    np.testing.assert_array_equal(fr_g.cpu().numpy(), fr_o)


def test_quadratic_eval_bitexact():
    rng = np.random.default_rng(5)
    for n, d, m in [(92, 12, 3), (257, 1024, 3), (31, 70, 10), (8, 1, 2)]:
        print(f"quadratic case n={n} d={d} m={m}", flush=True)
        x = rng.random((n, d), dtype=np.float32)
        c = rng.random((m, d), dtype=np.float32)
        f_o = O.eval_quadratic(x, c)
        f_g, cv_g = K.eval_quadratic(_f(x), _f(c))
        np.testing.assert_array_equal(f_g.cpu().numpy(), f_o)
        assert not cv_g.cpu().numpy().any()


def _draw_set(rng, n, d):
    pairs = n // 2
    return {
        "mate_a": rng.integers(0, n, n).astype(np.int32),
        "mate_b": rng.integers(0, n, n).astype(np.int32),
        "sbx_u": rng.random(pairs * d, dtype=np.float32),
        "sbx_p": rng.random(pairs * d, dtype=np.float32),
        "sbx_sign": rng.integers(0, 2, pairs * d).astype(np.int32),
        "sbx_copy": rng.random(pairs * d, dtype=np.float32),
        "pm_gate": rng.random(n * d, dtype=np.float32),
        "pm_delta": rng.random(n * d, dtype=np.float32),
    }


@pytest.mark.parametrize("n,d,copy_p", [(92, 12, 0.0), (64, 1024, 0.5), (31, 7, 0.0)])
def test_reproduce_injected(n, d, copy_p):
    rng = np.random.default_rng(n + d)
    x = rng.random((n, d), dtype=np.float32)
    lb, ub = -1.0, 2.0
    x = (lb + (ub - lb) * x).astype(np.float32)
    bounds = np.tile(np.array([lb, ub], np.float32), d)
    rank = rng.integers(0, 5, n).astype(np.int32)
    cv = np.where(rng.random(n) < 0.2, rng.random(n), 0.0).astype(np.float32)
    dr = _draw_set(rng, n, d)
    print(f"reproduce injected n={n} d={d}", flush=True)
    off_g, win_g = K.reproduce(_f(x), _f(bounds), rank=_f(rank), cv=_f(cv), sbx=(15.0, 0.9, copy_p),
                               pm=(20.0, 1.0), draws=dr)
    win_o = O.tournament(dr["mate_a"], dr["mate_b"], rank, cv)
    np.testing.assert_array_equal(win_g.cpu().numpy(), win_o)
    off_o = O.sbx(x, win_o, bounds, dr["sbx_u"], dr["sbx_p"], dr["sbx_sign"], dr["sbx_copy"], 15.0,
                  0.9, copy_p, n)
    off_o = O.pm(off_o, bounds, dr["pm_gate"], dr["pm_delta"], 20.0, 1.0)
    np.testing.assert_array_equal(off_g.cpu().numpy(), off_o)


def test_reproduce_generated_draws():
    n, d, seed, gen = 128, 33, 77, 5
    rng = np.random.default_rng(1)
    x = rng.random((n, d), dtype=np.float32)
    bounds = np.tile(np.array([0.0, 1.0], np.float32), d)
    rank = rng.integers(0, 3, n).astype(np.int32)
    cv = np.zeros(n, np.float32)
    print("reproduce generated draws", flush=True)
    off_g, _ = K.reproduce(_f(x), _f(bounds), rank=_f(rank), cv=_f(cv), seed=seed, generation=gen)
    ca, cb = O.draw_mating(seed, gen, n, n)
    win = O.tournament(ca, cb, rank, cv)
    ub, up, sg, uc = O.draw_sbx(seed, gen, n // 2, d)
    off = O.sbx(x, win, bounds, ub, up, sg, uc, 30.0, 1.0, 0.0, n)
    ug, ud = O.draw_pm(seed, gen, n, d)
    off = O.pm(off, bounds, ug, ud, 20.0, 1.0)
    np.testing.assert_array_equal(off_g.cpu().numpy(), off)


def test_init_population():
    bounds = np.tile(np.array([-2.0, 3.0], np.float32), 10)
    x_g = K.init_population(50, 10, _f(bounds), 99)
    np.testing.assert_array_equal(x_g.cpu().numpy(), O.draw_init(99, 50, 10, bounds))


def _select_case(rng, n, m, kind):
    if kind == "ties":
        fm = tie_objectives(rng, m, 2 * n, levels=5)
    elif kind == "sphere":
        fm = sphere_objectives(rng, m, 2 * n)
    else:
        fm = rng.random((m, 2 * n), dtype=np.float32)
    cvm = np.zeros(2 * n, np.float32)
    if kind == "constrained":
        cvm = np.where(rng.random(2 * n) < 0.4, rng.random(2 * n), 0.0).astype(np.float32)
    return fm, cvm


@pytest.mark.parametrize("n,m,kind,seed", [(46, 3, "rand", 0), (92, 3, "ties", 1), (200, 3, "sphere", 2),
                                           (128, 5, "rand", 3), (64, 8, "constrained", 4),
                                           (256, 10, "ties", 5), (1024, 3, "sphere", 6),
                                           (512, 2, "rand", 7)])
def test_nsga3_select(n, m, kind, seed):
    rng = np.random.default_rng(seed)
    fm, cvm = _select_case(rng, n, m, kind)
    d = 4
    xm = np.zeros((2 * n, d), np.float32)
    xm[:, 0] = np.arange(2 * n, dtype=np.float32)        # row identity
    xm[:, 1:] = rng.random((2 * n, d - 1), dtype=np.float32)
    w = O.das_dennis(n, m)
    ideal0 = np.minimum(fm[:, :n].min(axis=1), np.float32(0.05)).astype(np.float32)
    ua = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    ub = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    print(f"nsga3_select n={n} m={m} kind={kind} seed={seed}", flush=True)
    chosen, fr_mix, st, ideal_o = O.nsga3_select(fm, cvm, w, ideal0, ua, ub)
    par = K.Population(_f(xm[:n]), _f(fm[:, :n]), _f(cvm[:n]))
    ofs = K.Population(_f(xm[n:]), _f(fm[:, n:]), _f(cvm[n:]))
    ideal_g = _f(ideal0)
    nxt, nfr, stg = K.nsga3_select(par, ofs, _f(w), ideal_g, draws={"niche_a": ua, "niche_b": ub})
    stg = stg.cpu().numpy()
    assert stg[0] == st[0] and stg[1] == st[1] and stg[2] == st[2], (stg, st)
    got = nxt.x[:, 0].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(got, chosen)
    np.testing.assert_array_equal(nxt.f.cpu().numpy(), fm[:, chosen])
    np.testing.assert_array_equal(nxt.cv.cpu().numpy(), cvm[chosen])
    np.testing.assert_array_equal(nxt.x.cpu().numpy(), xm[chosen])
    np.testing.assert_array_equal(nfr.cpu().numpy(), fr_mix[chosen])
    np.testing.assert_array_equal(ideal_g.cpu().numpy(), ideal_o)


def test_nsga3_select_bench_full_size():
    """The standalone selection bench at its full size: 2N = 65536 rows s*u
    (u uniform on the positive unit-sphere orthant, s in [1, 2]), all
    feasible, K = 32640 Das-Dennis directions; fronts, the chosen set and
    its order against the oracle."""
    n, m = 32768, 3
    rng = np.random.default_rng(2024)
    fm = sphere_objectives(rng, m, 2 * n)
    cvm = np.zeros(2 * n, np.float32)
    xm = np.zeros((2 * n, 2), np.float32)
    xm[:, 0] = np.arange(2 * n, dtype=np.float32)
    w = O.das_dennis(n, m)
    assert w.shape == (3, 32640)
    ideal0 = fm[:, :n].min(axis=1).astype(np.float32)
    ua = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    ub = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    print("nsga3_select selection bench n=32768 m=3 K=32640", flush=True)
    chosen, fr_mix, st, ideal_o = O.nsga3_select(fm, cvm, w, ideal0, ua, ub)
    par = K.Population(_f(xm[:n]), _f(fm[:, :n]), _f(cvm[:n]))
    ofs = K.Population(_f(xm[n:]), _f(fm[:, n:]), _f(cvm[n:]))
    ideal_g = _f(ideal0)
    nxt, nfr, stg = K.nsga3_select(par, ofs, _f(w), ideal_g, draws={"niche_a": ua, "niche_b": ub})
    stg = stg.cpu().numpy()
    assert stg[0] == st[0] and stg[1] == st[1] and stg[2] == st[2], (stg, st)
    np.testing.assert_array_equal(nxt.x[:, 0].cpu().numpy().astype(np.int64), chosen)
    np.testing.assert_array_equal(nfr.cpu().numpy(), fr_mix[chosen])
    np.testing.assert_array_equal(nxt.f.cpu().numpy(), fm[:, chosen])


def test_nsga3_select_generated_draws():
    n, m, seed, gen = 300, 3, 2887, 4
    rng = np.random.default_rng(11)
    fm, cvm = _select_case(rng, n, m, "sphere")
    xm = np.zeros((2 * n, 2), np.float32)
    xm[:, 0] = np.arange(2 * n)
    w = O.das_dennis(n, m)
    ideal0 = fm[:, :n].min(axis=1).astype(np.float32)
    ua, ub = O.draw_niche(seed, gen, 2 * n)
    chosen, fr_mix, st, _ = O.nsga3_select(fm, cvm, w, ideal0, ua, ub)
    par = K.Population(_f(xm[:n]), _f(fm[:, :n]), _f(cvm[:n]))
    ofs = K.Population(_f(xm[n:]), _f(fm[:, n:]), _f(cvm[n:]))
    nxt, nfr, _ = K.nsga3_select(par, ofs, _f(w), _f(ideal0), seed=seed, generation=gen)
    np.testing.assert_array_equal(nxt.x[:, 0].cpu().numpy().astype(np.int64), chosen)


@pytest.mark.parametrize("n,active,gen", [(92, 92, 0), (120, 37, 3), (256, 1, 5), (1001, 640, 9)])
def test_random_mating_matches_oracle(n, active, gen):
    """RandomMating::select (mating_selection.cu:39-62): candidate t is
    draw % active over the live parents, with the live count read on the
    device (the RVEA path), bit for bit against the oracle's draw."""
    d = 4
    rng = np.random.default_rng(n + gen)
    x = torch.from_numpy(rng.random((n, d), dtype=np.float32)).to(DEV)
    bounds = torch.from_numpy(np.tile(np.array([0.0, 1.0], np.float32), d)).to(DEV)
    act = torch.tensor([active], dtype=torch.int32, device=DEV)
    print(f"random mating n={n} active={active} gen={gen}", flush=True)
    _, idx = K.reproduce(x, bounds, mating=1, active=n, active_device=act, seed=11, generation=gen)
    want = O.draw_mating(11, gen, n, active, pair=False)
    np.testing.assert_array_equal(idx.cpu().numpy(), want)
    assert idx.max().item() < active
