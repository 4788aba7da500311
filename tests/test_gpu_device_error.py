"""A front-peeling launch whose bounded barrier wait gives up must be reported,
never leave wrong fronts silently.  The host lowers the barrier's poll limit
(moea_debug_set_spin_limit) so the waits give up on their own while every
block still exits normally; each call then either reports the failure or, if
no wait happened to give up, returns exactly the normal result."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

import cuda_moea as cm  # noqa: E402
from cuda_moea import _lib  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


@pytest.fixture
def tiny_spin_limit():
    _lib.lib.moea_debug_set_spin_limit(1)
    try:
        yield
    finally:
        _lib.lib.moea_debug_set_spin_limit(0)


def _sphere(rng, m, n):
    g = np.abs(rng.standard_normal((n, m)))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    return torch.from_numpy(np.ascontiguousarray((rng.uniform(1, 2, (n, 1)) * u).T.astype(np.float32))).to(DEV)


def test_nds_rank_reports_a_gave_up_peel(tiny_spin_limit):
    rng = np.random.default_rng(5)
    f = _sphere(rng, 3, 65536)
    cv = torch.zeros(65536, dtype=torch.float32, device=DEV)
    _lib.lib.moea_debug_set_spin_limit(0)
    ref_fr, ref_mf = K.nds_rank(f, cv, 32768)
    _lib.lib.moea_debug_set_spin_limit(1)
    failed = 0
    for attempt in range(4):
        print(f"nds_rank attempt {attempt} with a 1-poll barrier limit", flush=True)
        try:
            fr, mf = K.nds_rank(f, cv, 32768)
        except _lib.MoeaError as e:
            assert "gave up" in str(e)
            failed += 1
            continue
        assert mf == ref_mf and torch.equal(fr, ref_fr)
    assert failed > 0
    _lib.lib.moea_debug_set_spin_limit(0)
    fr, mf = K.nds_rank(f, cv, 32768)             # the default limit works again
    assert mf == ref_mf and torch.equal(fr, ref_fr)


def test_nsga3_run_reports_a_gave_up_peel(tiny_spin_limit):
    def make():
        return cm.NSGA3(population_size=8192, max_generations=3, problem=cm.QuadraticCenters(16, 3, seed=3),
                        seed=1, print_progress=False)
    _lib.lib.moea_debug_set_spin_limit(0)
    ref = make().run()
    _lib.lib.moea_debug_set_spin_limit(1)
    failed = 0
    for attempt in range(3):
        print(f"NSGA-III run attempt {attempt} with a 1-poll barrier limit", flush=True)
        alg = make()
        try:
            res = alg.run()
        except RuntimeError as e:
            assert "gave up" in str(e)
            failed += 1
            # reset clears the flag: with the default limit the next run (a new
            # seed, as every reset draws one) finishes without an error
            _lib.lib.moea_debug_set_spin_limit(0)
            alg.reset()
            alg.run()
            _lib.lib.moea_debug_set_spin_limit(1)
            continue
        assert torch.equal(res.variables, ref.variables) and torch.equal(res.objectives, ref.objectives)
    assert failed > 0


def test_batch_reports_a_gave_up_peel(tiny_spin_limit):
    def runs():
        return [cm.NSGA3(population_size=n, max_generations=2, problem=cm.QuadraticCenters(12, 3, seed=i),
                         seed=i, print_progress=False) for i, n in enumerate((4096, 2048, 512))]
    _lib.lib.moea_debug_set_spin_limit(0)
    ref = cm.BatchNSGA3(runs()).run()
    _lib.lib.moea_debug_set_spin_limit(1)
    failed = 0
    for attempt in range(3):
        print(f"batch attempt {attempt} with a 1-poll barrier limit", flush=True)
        try:
            got = cm.BatchNSGA3(runs()).run()
        except RuntimeError as e:
            assert "gave up" in str(e)
            failed += 1
            continue
        for a, b in zip(got, ref):
            assert torch.equal(a.variables, b.variables) and torch.equal(a.objectives, b.objectives)
    assert failed > 0
