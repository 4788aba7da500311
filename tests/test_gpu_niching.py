"""Niching paths of moea_nsga3_select select the same survivors as the
oracle, bit for bit: one workgroup grouping the last front by direction (in
LDS or, past its capacity, in the workspace), the same workgroup's radix
sorts (LDS or workspace; also its fallback for a niche of more than 1024
members), and three device-wide sorts of all 2N slots."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import oracle as O  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
# (mode, LDS capacity): each path in LDS, the one-workgroup paths also with a
# capacity of one member (the rest in the workspace), the device-wide sorts
PATHS = [(1, 0), (1, 1), (3, 0), (3, 1), (2, 0)]


def _f(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _objectives(rng, kind, m, nm):
    if kind == "plane":                              # every row non-dominated: one front
        g = rng.random((m, nm)) + 1e-3
        return (g / g.sum(axis=0)).astype(np.float32)
    if kind == "clump":                              # one front; 40 % of it a single point
        g = rng.random((m, nm)) + 1e-3
        g[:, : (2 * nm) // 5] = g[:, :1]
        return (g / g.sum(axis=0)).astype(np.float32)
    if kind == "ties":
        return (rng.integers(0, 5, size=(m, nm)) / 5).astype(np.float32)
    g = np.abs(rng.standard_normal((nm, m)))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    return np.ascontiguousarray((rng.uniform(1.0, 2.0, (nm, 1)) * u).T.astype(np.float32))


@pytest.mark.parametrize("n,m,kind,seed", [(92, 3, "ties", 1), (300, 5, "sphere", 2), (1024, 3, "sphere", 3),
                                           (4000, 3, "plane", 4), (8192, 3, "plane", 5),
                                           (2048, 8, "ties", 6), (3000, 10, "sphere", 7),
                                           (700, 3, "clump", 8), (5000, 4, "clump", 9),
                                           (9000, 3, "sphere", 10), (4500, 3, "plane", 11)])
def test_niching_paths_equal_oracle(n, m, kind, seed):
    rng = np.random.default_rng(seed)
    fm = _objectives(rng, kind, m, 2 * n)
    cvm = np.zeros(2 * n, np.float32)
    xm = np.zeros((2 * n, 2), np.float32)
    xm[:, 0] = np.arange(2 * n, dtype=np.float32)
    w = O.das_dennis(n, m)
    ideal0 = fm[:, :n].min(axis=1).astype(np.float32)
    ua = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    ub = rng.integers(0, 2**32, 2 * n, dtype=np.uint64).astype(np.uint32)
    chosen, fr_mix, st, _ = O.nsga3_select(fm, cvm, w, ideal0, ua, ub)
    print(f"niching n={n} m={m} kind={kind}: n_prior={st[1]} n_last={st[2]}", flush=True)
    for mode, cap in PATHS:
        par = K.Population(_f(xm[:n]), _f(fm[:, :n]), _f(cvm[:n]))
        ofs = K.Population(_f(xm[n:]), _f(fm[:, n:]), _f(cvm[n:]))
        nxt, nfr, stg = K.nsga3_select(par, ofs, _f(w), _f(ideal0), draws={"niche_a": ua, "niche_b": ub},
                                       niching=(mode, cap))
        stg = stg.cpu().numpy()
        assert stg[1] == st[1] and stg[2] == st[2], (mode, cap, stg, st)
        np.testing.assert_array_equal(nxt.x[:, 0].cpu().numpy().astype(np.int64), chosen, err_msg=f"{mode} {cap}")
        np.testing.assert_array_equal(nfr.cpu().numpy(), fr_mix[chosen])


def test_niching_generated_draws_paths_agree():
    n, m, seed, gen = 2500, 3, 77, 9
    rng = np.random.default_rng(3)
    fm = _objectives(rng, "plane", m, 2 * n)
    cvm = np.zeros(2 * n, np.float32)
    xm = np.zeros((2 * n, 2), np.float32)
    xm[:, 0] = np.arange(2 * n)
    w = O.das_dennis(n, m)
    ideal0 = fm[:, :n].min(axis=1).astype(np.float32)
    ua, ub = O.draw_niche(seed, gen, 2 * n)
    chosen, _, _, _ = O.nsga3_select(fm, cvm, w, ideal0, ua, ub)
    for mode, cap in PATHS:
        par = K.Population(_f(xm[:n]), _f(fm[:, :n]), _f(cvm[:n]))
        ofs = K.Population(_f(xm[n:]), _f(fm[:, n:]), _f(cvm[n:]))
        nxt, _, _ = K.nsga3_select(par, ofs, _f(w), _f(ideal0), seed=seed, generation=gen, niching=(mode, cap))
        np.testing.assert_array_equal(nxt.x[:, 0].cpu().numpy().astype(np.int64), chosen, err_msg=f"{mode} {cap}")


def test_niching_rejects_bad_arguments():
    n, m = 46, 3
    rng = np.random.default_rng(0)
    fm = _objectives(rng, "sphere", m, 2 * n)
    z = np.zeros(2 * n, np.float32)
    x = np.zeros((2 * n, 2), np.float32)
    par = K.Population(_f(x[:n]), _f(fm[:, :n]), _f(z[:n]))
    ofs = K.Population(_f(x[n:]), _f(fm[:, n:]), _f(z[n:]))
    w = _f(O.das_dennis(n, m))
    for bad in ((4, 0), (1, -1), (-1, 0)):
        with pytest.raises(ValueError, match="niching"):
            K.nsga3_select(par, ofs, w, _f(fm[:, :n].min(axis=1)), niching=bad)
