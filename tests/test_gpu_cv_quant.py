"""NSGA-III CV quantization (selection/nsga3/kernels/cv_quantization.cu):
the bins against a float32 restatement, and a selection with quantization
against the oracle's selection of the same rows with their violations
replaced by those bins (only the sort sees the bins)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
import cuda_moea as cm  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def bins_f32(cv, bins, clip, log_alpha=0.0, eps=0.0):
    """cv_quantization.cu:8-34 in float32 (linear space; log space via log1p)."""
    cv = cv.astype(np.float32)
    u = np.minimum(np.maximum(cv, np.float32(0)), np.float32(clip))
    if log_alpha > 0:
        x = np.log1p(np.float32(log_alpha) * u).astype(np.float32) * np.float32(1.0 / np.log1p(np.float32(log_alpha * clip)))
    else:
        x = u * (np.float32(1.0) / np.float32(clip))
    v = (x * np.float32(bins)).astype(np.float32)
    b = np.clip(1 + np.floor(v).astype(np.int64), 1, bins)
    infeasible = (cv != 0) if eps == 0 else (cv > eps)
    return np.where(infeasible, b, 0).astype(np.float32), v


def _violations(rng, n):
    cv = np.exp(rng.standard_normal(n)).astype(np.float32)
    cv[rng.random(n) < 0.4] = 0.0
    cv[:5] = [0.0, -0.5, 1e-9, 3.0, 1e6]          # zero, negative, tiny, exact boundary, clipped
    return cv


@pytest.mark.parametrize("bins,clip,log_alpha,eps", [(8, 3.0, 0.0, 0.0), (5, 10.0, 2.0, 0.0),
                                                     (16, 1.0, 0.0, 0.05), (1, 2.0, 0.0, 0.0)])
def test_quantize_cv_bins(bins, clip, log_alpha, eps):
    rng = np.random.default_rng(bins)
    cv = _violations(rng, 5000)
    got = K.quantize_cv(torch.from_numpy(cv).to(DEV), bins=bins, clip_upper=clip, log_alpha=log_alpha,
                        feasibility_epsilon=eps).cpu().numpy()
    want, v = bins_f32(cv, bins, clip, log_alpha, eps)
    # the log space rests on log1pf, whose last bit may differ from the host's:
    # only values within rounding of a bin edge may land one bin apart
    edge = np.abs(v - np.round(v)) < (1e-4 if log_alpha > 0 else 0.0)
    assert ((got == want) | edge).all(), (got[got != want][:5], want[got != want][:5])
    assert np.isin(got, np.arange(bins + 1)).all()


@pytest.mark.parametrize("n,m,seed", [(500, 3, 1), (2000, 5, 2)])
def test_selection_with_quantization_equals_oracle_on_binned_violations(n, m, seed):
    rng = np.random.default_rng(seed)
    g = np.abs(rng.standard_normal((2 * n, m)))
    fm = np.ascontiguousarray((g / np.linalg.norm(g, axis=1, keepdims=True)).T.astype(np.float32))
    cvm = _violations(rng, 2 * n)
    xm = np.zeros((2 * n, 2), np.float32)
    xm[:, 0] = np.arange(2 * n)
    w = O.das_dennis(n, m)
    ideal0 = fm[:, :n].min(axis=1).astype(np.float32)
    ua = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    ub = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    q = dict(bins=6, clip_upper=4.0)
    binned, _ = bins_f32(cvm, 6, 4.0)
    chosen, fr_mix, _, _ = O.nsga3_select(fm, binned, w, ideal0, ua, ub)
    print(f"cv quant select n={n} m={m}", flush=True)

    def f(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    par = K.Population(f(xm[:n]), f(fm[:, :n]), f(cvm[:n]))
    ofs = K.Population(f(xm[n:]), f(fm[:, n:]), f(cvm[n:]))
    nxt, nfr, _ = K.nsga3_select(par, ofs, f(w), f(ideal0), draws={"niche_a": ua, "niche_b": ub}, cv_quant=q)
    rows = nxt.x[:, 0].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(rows, chosen)
    np.testing.assert_array_equal(nfr.cpu().numpy(), fr_mix[chosen])
    np.testing.assert_array_equal(nxt.cv.cpu().numpy(), cvm[chosen])    # raw violations survive


def test_run_with_quantization_runs_on_raw_violations():
    sel = cm.NSGA3EnvironmentSelector(cv_bins=4, cv_clip_upper=2.0)
    alg = cm.NSGA3(population_size=92, max_generations=15, problem=cm.C1DTLZ1(dimension=12, objectives=3),
                   environment_selector=sel, seed=1, print_progress=False)
    alg.initialize()
    cv0 = alg.population.constraints.cpu().numpy()
    assert (cv0 != np.round(cv0)).any()            # the initial violations are raw, not bins
    res = alg.run()
    assert (res.constraints.cpu().numpy() >= 0).all() and res.auxiliary_indices.shape == (92,)
