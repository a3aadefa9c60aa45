"""GPU: vox8_block_dots_noalloc (the allocation-free entry point, a fixed
library workspace instead of the context's) against vox8_block_dots on the
same inputs, through the C ABI, on two calls in a row, the second with fewer
vectors and a shorter length than the first (the workspace keeps the first
call's partials; nothing of them may leak into the second result).  Also:
k beyond the workspace limit is rejected with an error.  Bar: the same
fixed-order reductions, so bitwise equal; checked to 1e-14 relative."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ref(fem_ctx, n, k, X, ld, v):
    from vox8 import _lib
    out = (C.c_double * (2 * k + 1))()
    _lib.call("vox8_block_dots", fem_ctx, n, k, _lib.ptr(X), ld, _lib.ptr(v), out)
    return np.array(out[:])


def test_block_dots_noalloc_matches_context_entry(gpu):
    from vox8 import _lib
    ctx = _lib.create(8, 4, 4, torch.cuda.current_stream().cuda_stream)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(5)
    try:
        for n, k, ld in ((300_001, 12, 300_005), (70_000, 3, 70_000)):
            X = torch.tensor(rng.standard_normal((k, ld)), device=gpu)
            v = torch.tensor(rng.standard_normal(n), device=gpu)
            d_out = torch.full((2 * k + 1,), float("nan"), dtype=torch.float64, device=gpu)
            _lib.call("vox8_block_dots_noalloc", st, n, k, _lib.ptr(X), ld, _lib.ptr(v), _lib.ptr(d_out))
            got = d_out.cpu().numpy()
            want = _ref(ctx, n, k, X, ld, v)
            Xh, vh = X.cpu().numpy()[:, :n], v.cpu().numpy()
            exact = np.concatenate([Xh @ vh, (Xh * Xh).sum(axis=1), [vh @ vh]])
            assert np.all(np.isfinite(got))
            assert np.max(np.abs(got - want) / np.abs(want)) < 1e-14, (got, want)
            assert np.max(np.abs(got - exact) / np.abs(exact).max()) < 1e-12
        bad = _lib.load().vox8_block_dots_noalloc(st, 100, 17, _lib.ptr(X), 100, _lib.ptr(v),
                                                  _lib.ptr(d_out))
        assert bad != 0 and b"1..16" in _lib.load().vox8_last_error()
    finally:
        _lib.destroy(ctx)
