"""The allocation-free weighted dot (tnt_weighted_dot_async_noalloc) against
tnt_weighted_dot_async: the same chunking and kernels, so the results agree
bit for bit -- on two calls in a row, the second with a smaller input, and
with no growth of the context's device pool."""
import ctypes as C

import numpy as np
import pytest

from tenet_amd import _native as N
from tenet_amd.tensor import Runtime

pytestmark = pytest.mark.gpu


def _put(rt, arr):
    b = C.c_void_p()
    rt._check(N.hip().tnt_buf_alloc(rt._ctx, max(1, len(arr)), N.F64, C.byref(b)))
    a = np.ascontiguousarray(arr, dtype=np.float64)
    if len(a):
        rt._check(N.hip().tnt_buf_upload(rt._ctx, b, a.ctypes.data, len(a)))
    return b


def _get(rt, b, n):
    out = np.empty(n)
    rt._check(N.hip().tnt_buf_download(rt._ctx, b, out.ctypes.data, n))
    return out


def _regions(rng, n, k):
    cuts = np.sort(rng.choice(np.arange(1, n), size=k - 1, replace=False))
    e = np.concatenate([[0], cuts, [n]])
    regs = (N.Region * k)(*[N.Region(int(a), int(b - a)) for a, b in zip(e[:-1], e[1:])])
    w = rng.uniform(1, 9, size=k)
    return regs, (C.c_double * k)(*w), e, w


def test_noalloc_dot_matches_async_on_two_calls():
    rt = Runtime(0)
    h = N.hip()
    try:
        rng = np.random.default_rng(17)
        n = 3_000_000  # 46 chunks of 65536 over 21 regions, then a smaller call
        x, y = rng.normal(size=n), rng.normal(size=n)
        bx, by, res = _put(rt, x), _put(rt, y), _put(rt, np.zeros(8))
        pool0 = C.c_size_t()
        for call, (size, k) in enumerate(((n, 21), (70_000, 5))):
            regs, w, e, wv = _regions(rng, size, k)
            rt._check(h.tnt_weighted_dot_async(rt._ctx, bx, by, regs, w, k, res, 2 * call))
            rt.synchronize()
            rt._check(h.tnt_ctx_pool_bytes(rt._ctx, C.byref(pool0)))
            rt._check(h.tnt_weighted_dot_async_noalloc(rt._ctx, bx, by, regs, w, k, res, 2 * call + 1))
            rt.synchronize()
            pool1 = C.c_size_t()
            rt._check(h.tnt_ctx_pool_bytes(rt._ctx, C.byref(pool1)))
            assert pool1.value == pool0.value  # nothing allocated
            got = _get(rt, res, 8)
            assert got[2 * call].tobytes() == got[2 * call + 1].tobytes(), (call, got[2 * call], got[2 * call + 1])
            want = sum(wi * float(x[a:b] @ y[a:b]) for wi, a, b in zip(wv, e[:-1], e[1:]))
            assert abs(got[2 * call + 1] - want) <= 1e-11 * abs(want) + 1e-12
        # beyond the scratch: a clear error, nothing queued
        big = (N.Region * 1)(N.Region(0, n))
        one = (C.c_double * 1)(1.0)
        many = 40000
        regs = (N.Region * many)(*[N.Region(i, 1) for i in range(many)])
        ws = (C.c_double * many)(*([1.0] * many))
        rc = h.tnt_weighted_dot_async_noalloc(rt._ctx, bx, by, regs, ws, many, res, 0)
        assert rc == N.E_UNSUPPORTED and b"scratch" in h.tnt_last_error(rt._ctx)
        # an empty region list writes 0
        rt._check(h.tnt_weighted_dot_async_noalloc(rt._ctx, bx, by, big, one, 0, res, 7))
        rt.synchronize()
        assert _get(rt, res, 8)[7] == 0.0
        for b in (bx, by, res):
            h.tnt_buf_free(rt._ctx, b)
    finally:
        rt.close()
