"""Built-in objective kernels against float64 references of the same
formulas: quadratic-center f_k = sum_j (x_j - c_kj)^2 and the seeded MLP
f_k = (W2 tanh(W1 x))_k^2 + 0.01 ||x - 0.5||^2.  Both are float32 reductions
whose summation order differs from the reference's, so they are compared with
a tolerance stated per test; whole runs on them are compared bit for bit
elsewhere (single run against batched and served runs)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


@pytest.mark.parametrize("n,d,m", [(1, 1, 1), (37, 12, 3), (300, 1000, 5), (64, 4096, 10)])
def test_quadratic_against_fp64(n, d, m):
    rng = np.random.default_rng(n + d + m)
    x = rng.random((n, d), dtype=np.float32)
    c = rng.random((m, d), dtype=np.float32)
    f, cv = K.eval_quadratic(torch.from_numpy(x).to(DEV), torch.from_numpy(c).to(DEV))
    ref = ((x.astype(np.float64)[None, :, :] - c.astype(np.float64)[:, None, :]) ** 2).sum(-1)
    got = f.cpu().numpy().astype(np.float64)
    # float32 sum of d terms in [0, 1]: relative error within a few d * 2^-24
    np.testing.assert_allclose(got, ref, rtol=4 * d * 2.0 ** -24 + 1e-7, atol=1e-6)
    assert np.all(cv.cpu().numpy() == 0.0)


@pytest.mark.parametrize("n,d,m,hidden", [(1, 3, 1, 1), (70, 40, 3, 256), (129, 300, 5, 100),
                                          (64, 2048, 8, 64), (33, 17, 10, 257)])
def test_mlp_against_fp64(n, d, m, hidden):
    rng = np.random.default_rng(n * 7 + d)
    prob = cm.SeededMLP(d, m, hidden=hidden, seed=n)
    x = rng.random((n, d), dtype=np.float32)
    f, cv = K.eval_mlp(torch.from_numpy(x).to(DEV), prob.w1.to(DEV), prob.w2.to(DEV))
    x64 = x.astype(np.float64)
    h = np.tanh(x64 @ prob.w1.double().numpy().T)
    y = h @ prob.w2.double().numpy().T
    ref = (y ** 2 + 0.01 * ((x64 - 0.5) ** 2).sum(1, keepdims=True)).T
    got = f.cpu().numpy().astype(np.float64)
    # W1 x is a float32 dot product of d terms of size ~1/sqrt(d): absolute error
    # ~ sqrt(d) 2^-24 per hidden unit, carried through tanh and W2 into y and y^2
    tol = 64 * np.sqrt(d) * 2.0 ** -24 * (1.0 + np.abs(ref))
    assert np.all(np.abs(got - ref) <= tol), float(np.max(np.abs(got - ref) / (1 + np.abs(ref))))
    assert got.shape == (m, n)
    assert np.all(cv.cpu().numpy() == 0.0)


def test_mlp_run_native():
    prob = cm.SeededMLP(64, 3, hidden=128, seed=3)
    res = cm.NSGA3(population_size=92, max_generations=5, problem=prob, seed=1,
                   print_progress=False).run()
    f = res.objectives.cpu().numpy()
    assert f.shape == (92, 3) and np.all(np.isfinite(f)) and np.all(f >= 0)
    assert (res.auxiliary_indices.cpu().numpy() == 0).any()
    # objectives of the returned population are the MLP of its variables
    x = res.variables.cpu().double().numpy()
    y = np.tanh(x @ prob.w1.double().numpy().T) @ prob.w2.double().numpy().T
    ref = y ** 2 + 0.01 * ((x - 0.5) ** 2).sum(1, keepdims=True)
    np.testing.assert_allclose(f, ref, rtol=1e-4, atol=1e-6)
