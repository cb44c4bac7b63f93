"""The allocation-free pair selection (tk_select_pairs_d_noalloc) against
tk_select_pairs_d, through the C ABI: bit-identical flags and counts on two
calls in a row (the second smaller, so stale scratch from the first must not
leak into it), no device memory taken by the library, and a clear error past
its 2^24-pair limit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stats(api, rng, n):
    st = np.zeros(2 * n, api.READ_STAT_DTYPE)
    nwin = rng.integers(1, 127, 2 * n).astype(np.uint32)
    nwin[rng.random(2 * n) < 0.04] = 0
    med = rng.integers(1, 400, 2 * n).astype(np.uint32)
    med[nwin == 0] = 0
    st["nwin"] = nwin
    st["median"] = med
    st["sum"] = nwin.astype(np.uint64) * np.maximum(med, 1).astype(np.uint64)
    st["sum"][nwin == 0] = 0
    st["sumsq_lo"] = st["sum"] * np.maximum(med, 1).astype(np.uint64)
    return st


def _run(api, ctx, fn, d_st, n, d_sel, min_cov):
    ns = C.c_uint64()
    rc = getattr(api.lib(), fn)(ctx.h, C.c_void_p(d_st.ptr), C.c_void_p(d_st.ptr + n * 32), n, 50, min_cov, 10000,
                                126, C.c_void_p(d_sel.ptr), C.byref(ns))
    api._check(rc, fn)
    ctx.sync()
    sel = np.zeros(n, np.uint8)
    d_sel.download(sel)
    return sel, ns.value


def test_noalloc_select_equals_allocating_select(ctx):
    from trinityrnaseq_amd import api
    rng = np.random.default_rng(2024)
    for n, min_cov in ((3_000_000, 0), (250_001, 2)):
        print("noalloc select pairs", n, "min_cov", min_cov, flush=True)
        st = _stats(api, rng, n)
        d_st = api.DeviceBuffer(ctx, st.nbytes)
        d_st.upload(st)
        d_a = api.DeviceBuffer(ctx, n + 64)
        d_b = api.DeviceBuffer(ctx, n + 64)
        want, nw = _run(api, ctx, "tk_select_pairs_d", d_st, n, d_a, min_cov)
        before = api.device_bytes_in_use()
        got, ng = _run(api, ctx, "tk_select_pairs_d_noalloc", d_st, n, d_b, min_cov)
        assert api.device_bytes_in_use() == before
        assert ng == nw and np.array_equal(got, want)
        assert 0 < nw < n
        for b in (d_st, d_a, d_b):
            b.free()


def test_noalloc_select_refuses_more_than_its_scratch(ctx):
    from trinityrnaseq_amd import api
    ns = C.c_uint64()
    rc = api.lib().tk_select_pairs_d_noalloc(ctx.h, C.c_void_p(256), C.c_void_p(512), (1 << 24) + 1, 50, 0, 10000,
                                             126, C.c_void_p(1024), C.byref(ns))
    assert rc == -1
    assert b"exceed" in api.lib().tk_last_error()
