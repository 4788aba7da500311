"""Native DTLZ1 / DTLZ2 objectives (the reference's built-in kernels,
problem/dtlz/dtlz_kernels.cu:28-190) against an FP64 restatement, at the
reference's native-objective tolerance (run_benchmark.py:54-61: rtol 2e-5,
atol 2e-4)."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402
from cuda_moea import _lib, kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def dtlz_fp64(kind, x, m):
    x = x.astype(np.float64)
    xm, xk = x[:, : m - 1], x[:, m - 1:]
    if kind == 1:
        g = 100.0 * (xk.shape[1] + ((xk - 0.5) ** 2 - np.cos(20 * math.pi * (xk - 0.5))).sum(1))
        base = 0.5 * (1 + g)
    else:
        base = 1.0 + ((xk - 0.5) ** 2).sum(1)
    f = np.empty((m, x.shape[0]))
    for i in range(m):
        v = base.copy()
        for j in range(m - 1 - i):
            v *= np.clip(xm[:, j], 0, 1) if kind == 1 else np.maximum(0.0, np.cos(xm[:, j] * math.pi / 2))
        if i > 0:
            xi = xm[:, m - 1 - i]
            v *= np.maximum(0.0, 1 - xi) if kind == 1 else np.sin(xi * math.pi / 2)
        f[i] = v
    return f


@pytest.mark.parametrize("kind,n,d,m", [(1, 300, 12, 3), (1, 1000, 500, 3), (1, 77, 40, 8),
                                        (2, 300, 12, 3), (2, 513, 300, 5), (2, 64, 10, 10)])
def test_dtlz_kernel_against_fp64(kind, n, d, m):
    rng = np.random.default_rng(kind * 100 + m)
    x = rng.random((n, d), dtype=np.float32)
    print(f"dtlz{kind} n={n} d={d} m={m}", flush=True)
    f, cv = K.eval_dtlz(kind, torch.from_numpy(x).to(DEV), m)
    np.testing.assert_allclose(f.cpu().numpy(), dtlz_fp64(kind, x, m), rtol=2e-5, atol=2e-4)
    assert not cv.cpu().numpy().any()


@pytest.mark.parametrize("cls,kind", [(cm.DTLZ1, 1), (cm.DTLZ2, 2)])
def test_dtlz_run_native(cls, kind):
    alg = cm.NSGA3(population_size=92, max_generations=20, problem=cls(dimension=12, objectives=3),
                   seed=3, print_progress=False)
    alg.initialize()
    assert alg._args.problem.kind == (_lib.PROBLEM_DTLZ1 if kind == 1 else _lib.PROBLEM_DTLZ2)
    res = alg.run()
    x = res.variables.cpu().numpy()
    np.testing.assert_allclose(res.objectives.cpu().numpy().T, dtlz_fp64(kind, x, 3), rtol=2e-5, atol=2e-4)
