"""The allocation-free sort (moea_nds_rank_static, workspace in a device array
inside the library) against moea_nds_rank with a caller workspace, through
the C ABI: bit-identical fronts and max_front on two calls in a row, the
second smaller, then a size beyond the array's limit is refused."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from cuda_moea import _lib  # noqa: E402
from cuda_moea import kernels as K  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _objectives(rng, m, n, ties):
    if ties:
        return rng.integers(0, 7, size=(m, n)).astype(np.float32) / np.float32(7)
    g = np.abs(rng.standard_normal((n, m)))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    return np.ascontiguousarray((rng.uniform(1.0, 2.0, (n, 1)) * u).T.astype(np.float32))


def _static(f, cv, stop):
    m, n = f.shape
    fronts = torch.full((n,), -7, dtype=torch.int32, device=DEV)
    mf = torch.full((1,), -7, dtype=torch.int32, device=DEV)
    st = torch.cuda.current_stream(DEV).cuda_stream
    _lib.check(_lib.lib.moea_nds_rank_static(f.data_ptr(), n, cv.data_ptr(), n, m, stop, fronts.data_ptr(),
                                             mf.data_ptr(), st))
    return fronts, int(mf.item())


def test_static_matches_workspace_path_twice_then_smaller():
    assert _lib.lib.moea_nds_static_max_n() == 8192
    rng = np.random.default_rng(71)
    # first call at the largest size, second smaller: stale state of the first
    # call must not leak into the second
    for n, m, ties, stop in ((8192, 3, False, 4096), (1500, 5, True, 1500), (300, 2, True, 97)):
        print(f"static nds case n={n} m={m} ties={ties} stop={stop}", flush=True)
        f = torch.from_numpy(_objectives(rng, m, n, ties)).to(DEV)
        cv = np.where(rng.random(n) < 0.1, rng.integers(1, 4, n), 0).astype(np.float32)
        cv = torch.from_numpy(cv).to(DEV)
        ref_fr, ref_mf = K.nds_rank(f, cv, stop)
        got_fr, got_mf = _static(f, cv, stop)
        torch.cuda.synchronize()
        assert got_mf == ref_mf
        assert torch.equal(got_fr, ref_fr)


def test_static_rejects_beyond_limit():
    n = _lib.lib.moea_nds_static_max_n() + 2
    f = torch.zeros((3, n), dtype=torch.float32, device=DEV)
    cv = torch.zeros(n, dtype=torch.float32, device=DEV)
    with pytest.raises(ValueError, match="static workspace"):
        _static(f, cv, n)
