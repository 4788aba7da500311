"""GPU parity of the RVEA path and of the MLP objective."""
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


@pytest.mark.parametrize("k_req,m", [(91, 3), (210, 5), (120, 8), (1000, 3), (300, 10)])
def test_gamma(k_req, m):
    v = O.das_dennis(k_req, m, kind=2)
    print(f"gamma k_req={k_req} m={m}", flush=True)
    np.testing.assert_array_equal(K.rvea_gamma(_f(v)).cpu().numpy(), O.rvea_gamma(v))


@pytest.mark.parametrize("n,m,active,constrained,seed", [(60, 3, 60, False, 0), (100, 5, 37, True, 1),
                                                          (210, 5, 210, False, 2), (64, 8, 20, False, 3),
                                                          (500, 3, 400, True, 4)])
def test_rvea_select(n, m, active, constrained, seed):
    rng = np.random.default_rng(seed)
    v = O.das_dennis(n, m, kind=2)
    gam = O.rvea_gamma(v)
    fm = (rng.random((m, 2 * n), dtype=np.float32) + 0.05).astype(np.float32)
    cvm = np.zeros(2 * n, np.float32)
    if constrained:
        cvm = np.where(rng.random(2 * n) < 0.3, rng.random(2 * n) * 2, 0.0).astype(np.float32)
    fm[:, active:n] = 1e15          # dead parent rows as the previous selection left them
    cvm[active:n] = 1e15
    pen = O.rvea_penalty(m, 7, 100, 2.0)
    xm = np.zeros((2 * n, 3), np.float32)
    xm[:, 0] = np.arange(2 * n)
    print(f"rvea_select n={n} m={m} active={active} seed={seed}", flush=True)
    sel, na, ref, th, apd = O.rvea_select(fm, cvm, v, gam, active, pen)
    par = K.Population(_f(xm[:n]), _f(fm[:, :n]), _f(cvm[:n]))
    ofs = K.Population(_f(xm[n:]), _f(fm[:, n:]), _f(cvm[n:]))
    nxt, na_g = K.rvea_select(par, ofs, _f(v), _f(gam), active, pen)
    assert na_g == na
    got = nxt.x[:na, 0].cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(got, sel)
    np.testing.assert_array_equal(nxt.f[:, :na].cpu().numpy(), fm[:, sel])
    if na < n:
        assert np.isnan(nxt.x[na:].cpu().numpy()).all()
        assert (nxt.f[:, na:].cpu().numpy() == np.float32(1e15)).all()
        assert (nxt.cv[na:].cpu().numpy() == np.float32(1e15)).all()


def test_rvea_adapt():
    v0 = O.das_dennis(300, 3, kind=2)
    rng = np.random.default_rng(9)
    f = (rng.random((3, 300), dtype=np.float32) * np.array([[1.0], [5.0], [0.2]], np.float32))
    out_o = O.rvea_adapt(v0, f, 250)
    out_g = K.rvea_adapt(_f(v0), _f(f), 250)
    np.testing.assert_array_equal(out_g.cpu().numpy(), out_o)


def test_mlp_objective_against_fp64():
    rng = np.random.default_rng(4)
    n, d, hidden, m = 300, 700, 256, 5
    x = rng.random((n, d), dtype=np.float32)
    w1 = (rng.standard_normal((hidden, d)) / np.sqrt(d)).astype(np.float32)
    w2 = (rng.standard_normal((m, hidden)) / np.sqrt(hidden)).astype(np.float32)
    f_g, cv_g = K.eval_mlp(_f(x), _f(w1), _f(w2))
    X = x.astype(np.float64)
    y = w2.astype(np.float64) @ np.tanh(w1.astype(np.float64) @ X.T)
    ref = y ** 2 + 0.01 * ((X - 0.5) ** 2).sum(1)[None, :]
    # the reference's native-objective check (run_benchmark.py:54-61): rtol 2e-5, atol 2e-4
    np.testing.assert_allclose(f_g.cpu().numpy(), ref, rtol=2e-5, atol=2e-4)
    assert not cv_g.cpu().numpy().any()
