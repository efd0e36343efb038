"""gdl_sort_unique_u64_noalloc (include/gdl_hip.h): the allocation-free sort +
unique gives exactly what gdl_sort_unique_u64 gives, through the C ABI, on
consecutive calls that reuse the fixed scratch array (the second smaller), and
makes no allocation through the context's allocator."""
import ctypes

import numpy as np
import pytest

from conftest import gdl

MAX_KEYS = 1 << 20


def _mem(ctx):
    m = ctx.mem_info()
    return m["in_use"], m["peak"], m["cached"]


def _noalloc(ctx, d_in, n, d_out, hi=-1, lo=-1):
    cnt = ctypes.c_uint64()
    st = ctx.eng.gdl_sort_unique_u64_noalloc(ctx.h, d_in, n, hi, lo, d_out, ctypes.byref(cnt))
    return st, cnt.value


def _with_alloc(ctx, d_in, n, d_out, hi=-1, lo=-1):
    cnt = ctypes.c_uint64()
    st = ctx.eng.gdl_sort_unique_u64(ctx.h, d_in, n, hi, lo, d_out, ctypes.byref(cnt))
    return st, cnt.value


@pytest.mark.gpu
def test_noalloc_sort_unique_matches_the_allocating_entry_point():
    rng = np.random.default_rng(11)
    big = rng.integers(0, 1 << 40, MAX_KEYS, dtype=np.uint64)
    big[::7] = big[3]  # duplicates
    small = rng.integers(0, 1 << 20, 5000, dtype=np.uint64) << np.uint64(32)
    tiny = rng.integers(0, 300, 900, dtype=np.uint64)  # one-block path (<= 4096 keys)
    with gdl.Context(device=0) as ctx:
        bufs = []
        try:
            cases = []
            for keys in (big, small, tiny):
                d = ctx.to_device(keys)
                o1, o2 = ctx.alloc(keys.size), ctx.alloc(keys.size)
                bufs += [d, o1, o2]
                cases.append((keys, d, o1, o2))
            for keys, d, o1, o2 in cases:  # big, then smaller inputs on the same scratch
                before = _mem(ctx)
                st, n_na = _noalloc(ctx, d.ptr, d.n, o1.ptr)
                assert st == 0, ctx.eng.gdl_last_error(ctx.h)
                assert _mem(ctx) == before  # nothing from the allocator
                st, n_ref = _with_alloc(ctx, d.ptr, d.n, o2.ptr)
                assert st == 0
                assert n_na == n_ref == np.unique(keys).size
                assert np.array_equal(o1.numpy(n_na), o2.numpy(n_ref))
                assert np.array_equal(o1.numpy(n_na), np.unique(keys))
            # stated bit bounds take the same path as measured ones
            keys, d, o1, o2 = cases[1]
            st, n_na = _noalloc(ctx, d.ptr, d.n, o1.ptr, hi=20, lo=0)
            assert st == 0 and np.array_equal(o1.numpy(n_na), np.unique(keys))
            # one key past the limit: refused, nothing written
            st, _ = _noalloc(ctx, d.ptr, MAX_KEYS + 1, o1.ptr)
            assert st == gdl.GDL_ERR_INVALID_ARG
            assert b"scratch limit" in ctx.eng.gdl_last_error(ctx.h)
            st, n0 = _noalloc(ctx, None, 0, None)
            assert st == 0 and n0 == 0
        finally:
            for b in bufs:
                b.free()
