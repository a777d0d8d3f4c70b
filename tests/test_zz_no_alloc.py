"""psr_isovalue_noalloc (the allocation-free isovalue reduction) through the C
ABI: the same bits as psr_isovalue on two calls in a row, the second with a
smaller input (the shared partials array keeps the first call's contents),
and sizes beyond its maximum rejected before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(count, dtype, seed):
    rng = np.random.default_rng(seed)
    values = rng.normal(size=count).astype(dtype)
    weights = np.exp2(rng.uniform(-20, 20, size=count)).astype(dtype)
    weights[::13] = 0
    weights[0] = 1
    return values, weights


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_isovalue_matches_pooled(cuda, dtype):
    import torch
    from warp_psr_amd import _lib
    lib = _lib.load()
    code = 1 if dtype == np.float64 else 0
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    for count, seed in ((3_000_001, 1), (70_001, 2), (1_000, 3)):
        print(f"noalloc isovalue {np.dtype(dtype).name} count {count}", flush=True)
        values, weights = _inputs(count, dtype, seed)
        v, w = torch.from_numpy(values).to(cuda), torch.from_numpy(weights).to(cuda)
        out = torch.full((2,), float("nan"), dtype=v.dtype, device=cuda)
        _lib.check(lib.psr_isovalue(code, v.data_ptr(), w.data_ptr(), count, out[0:].data_ptr(), stream))
        _lib.check(lib.psr_isovalue_noalloc(code, v.data_ptr(), w.data_ptr(), count, out[1:].data_ptr(), stream))
        a, b = out.cpu().numpy()
        assert np.isfinite(a)
        assert a.tobytes() == b.tobytes()


def test_noalloc_isovalue_rejects_oversized_input(cuda):
    import torch
    from warp_psr_amd import _lib
    lib = _lib.load()
    out = torch.empty(1, dtype=torch.float32, device=cuda)
    stream = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    # rejected from the size alone: the (dummy, non-null) inputs are never read
    rc = lib.psr_isovalue_noalloc(0, out.data_ptr(), out.data_ptr(), (1 << 26) + 1, out.data_ptr(), stream)
    assert rc == 2
    assert "at most" in lib.psr_last_error().decode()
