"""The allocation-free overlap (derek_hip_overlap_noalloc: device buffers,
split-K partial sums in a file-scope device array) against the allocating
entry point derek_hip_overlap on the same inputs, through the C-ABI: two
calls in a row, the second on a smaller input, each within 1e-13 (relative
to the largest element) of the allocating call; FP64 sums over O(npw)
products, whose order changes only when the row chunks are capped.  Device
memory comes from torch."""
import ctypes as C

import numpy as np
import pytest
import torch

from derek_amd import lib

pytestmark = pytest.mark.gpu


def _p(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


def _noalloc(ctx, X, Y, herm):
    """S (m x n) from device copies of X (m, npw) and Y (n, npw)."""
    m, npw = X.shape
    n = m if herm else Y.shape[0]
    dX = torch.from_numpy(np.ascontiguousarray(X)).to("cuda")
    dY = dX if herm else torch.from_numpy(np.ascontiguousarray(Y)).to("cuda")
    dS = torch.full((n, m), complex(np.nan, np.nan), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    lib.check(ctx.lib.derek_hip_overlap_noalloc(ctx.h, C.c_int64(npw), C.c_int32(m), C.c_int32(n), _p(dX), _p(dY),
                                                _p(dS)))
    ctx.sync()
    return dS.cpu().numpy().T.copy()


def test_overlap_noalloc_matches_overlap():
    print("overlap_noalloc", flush=True)
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1)
    rng = np.random.default_rng(21)
    # first a large Hermitian product (its row chunks are capped to fit the
    # array), then a smaller general one
    for npw, m, n, herm in [(62669, 160, 160, True), (5003, 24, 40, False)]:
        print("case", npw, m, n, herm, flush=True)
        X = rng.normal(size=(m, npw)) + 1j * rng.normal(size=(m, npw))
        Y = X if herm else rng.normal(size=(n, npw)) + 1j * rng.normal(size=(n, npw))
        ref = ctx.overlap(X, X if herm else Y)
        got = _noalloc(ctx, X, Y, herm)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"npw {npw} {m}x{n} herm {herm}: max rel diff {err:.2e}")
        assert got.shape == (m, n)
        assert np.all(np.isfinite(got))
        assert err < 1e-13
    ctx.close()


def test_overlap_noalloc_refusals():
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1)
    d = torch.zeros(16, dtype=torch.complex128, device="cuda")
    # m * n beyond the scratch capacity
    with pytest.raises(lib.DerekError) as e:
        lib.check(ctx.lib.derek_hip_overlap_noalloc(ctx.h, C.c_int64(1), C.c_int32(2048), C.c_int32(2048), _p(d),
                                                    _p(d), _p(d)))
    assert e.value.code == lib.DEREK_ERR_ARG and "scratch" in str(e.value)
    ctx.close()
    mx = lib.Context((8, 8, 8), 1, 1, 1.0, 1, lib.DEREK_MIXED)
    with pytest.raises(lib.DerekError) as e:
        lib.check(mx.lib.derek_hip_overlap_noalloc(mx.h, C.c_int64(4), C.c_int32(2), C.c_int32(2), _p(d), _p(d),
                                                   _p(d)))
    assert e.value.code == lib.DEREK_ERR_ARG
    mx.close()
