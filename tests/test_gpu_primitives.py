"""Kernel-level parity of the HIP engine through the C ABI (bit-exact).

References are textbook numpy routines (np.sort / np.unique / set algebra /
searchsorted), restated independently of both the engine and the oracle.
"""
import numpy as np
import pytest

from conftest import gdl

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)
M32 = np.uint64(0xffffffff)


def rand_keys(n, hi_bits=32, lo_bits=32, rng=RNG):
    hi = rng.integers(0, 1 << hi_bits, n, dtype=np.uint64) if hi_bits else np.zeros(n, np.uint64)
    lo = rng.integers(0, 1 << lo_bits, n, dtype=np.uint64) if lo_bits else np.zeros(n, np.uint64)
    return (hi << np.uint64(32)) | lo


SIZES = [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 100_000, 1_000_003]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("bits", [(32, 32), (18, 18), (15, 15), (1, 0), (0, 7),
                                  (9, 9), (13, 14), (17, 17), (27, 0)])
def test_sort_matches_numpy(ctx, n, bits):
    # 8-bit digits, and 9-bit digits where they save a pass (18, 27, 34 and
    # 36 key bits: 2, 3, 4 and 4 passes)
    k = rand_keys(n, *bits)
    got = ctx.sort(k, *bits)
    assert np.array_equal(got, np.sort(k))


@pytest.mark.parametrize("n", SIZES)
def test_sort_unique_matches_numpy(ctx, n):
    k = rand_keys(n, 12, 12)  # many duplicates
    assert np.array_equal(ctx.sort_unique(k, 12, 12), np.unique(k))
    assert np.array_equal(ctx.sort_unique(k, -1, -1), np.unique(k))  # measured bounds


def test_sort_adversarial(ctx):
    n = 300_000
    cases = [np.full(n, 12345, np.uint64), np.arange(n, dtype=np.uint64),
             np.arange(n, dtype=np.uint64)[::-1].copy(),
             gdl.pack(np.full(n, 0xffffffff), np.arange(n)[::-1] % 7 + 0xfffffff0)]
    for k in cases:
        assert np.array_equal(ctx.sort(k, 32, 32), np.sort(k))
        assert np.array_equal(ctx.sort_unique(k, 32, 32), np.unique(k))


def test_key_bits(ctx):
    k = gdl.pack([3, 200000, 7], [1, 2, 131071])
    assert ctx.key_bits(k) == (18, 17)


@pytest.mark.parametrize("n", [0, 1, 5000, 1_000_000])
def test_unique_matches_numpy(ctx, n):
    k = np.sort(rand_keys(n, 8, 8))
    assert np.array_equal(ctx.unique(k), np.unique(k))


@pytest.mark.parametrize("na,nb", [(0, 0), (0, 10), (10, 0), (1, 1), (3000, 2000),
                                   (1_000_000, 300_000), (50, 2_000_000)])
def test_merge_union_difference_match_set_algebra(ctx, na, nb):
    a = np.unique(rand_keys(na, 11, 11))
    b = np.unique(rand_keys(nb, 11, 11))
    u = np.union1d(a, b)
    d = np.setdiff1d(b, a)
    assert np.array_equal(ctx.merge_union(a, b), u)
    assert np.array_equal(ctx.set_difference(b, a), d)
    gu, gd = ctx.merge_diff(a, b)
    assert np.array_equal(gu, u) and np.array_equal(gd, d)


def test_merge_diff_identical_and_disjoint(ctx):
    a = np.unique(rand_keys(500_000, 20, 20))
    gu, gd = ctx.merge_diff(a, a)
    assert np.array_equal(gu, a) and gd.size == 0
    b = a + np.uint64(1 << 63)
    gu, gd = ctx.merge_diff(a, b)
    assert np.array_equal(gu, np.concatenate([a, b])) and np.array_equal(gd, b)


def np_runs(sorted_keys, prefixes, shift):
    p = sorted_keys >> np.uint64(shift)
    lo = np.searchsorted(p, prefixes, "left")
    hi = np.searchsorted(p, prefixes, "right")
    return lo, hi - lo


@pytest.mark.parametrize("prefix_cols,layout", [(1, "dense"), (1, "holes"), (1, "spread"),
                                               (2, "dense")])
@pytest.mark.parametrize("lf", [0.5, 0.8])
def test_index_lookup_runs(ctx, prefix_cols, layout, lf):
    # one column of small ids gets the direct-mapped index; ids spread over
    # 32 bits (or a two-column prefix) get the hash index
    ctx.eng.gdl_set_load_factor(ctx.h, lf)
    keys = rand_keys(200_000, 14, 16)
    if layout == "holes":  # even ids only: odd ids inside the range are absent
        keys = gdl.pack((keys >> np.uint64(32)) & ~np.uint64(1), keys & M32)
    elif layout == "spread":
        keys = gdl.pack((keys >> np.uint64(32)) * np.uint64(100_003), keys & M32)
    keys = np.unique(keys)
    shift = 32 if prefix_cols == 1 else 0
    d = ctx.to_device(keys)
    idx = ctx.index_build(d, prefix_cols)
    assert idx.dense == (1 if prefix_cols == 1 and layout != "spread" else 0)
    present = np.unique(keys >> np.uint64(shift))
    assert idx.n_distinct == present.size
    if prefix_cols == 2:
        absent = present[:1000] + np.uint64(1 << 40)
    elif layout == "holes":
        absent = np.arange(1, 2001, 2, dtype=np.uint64)
    elif layout == "spread":
        absent = present[:1000] + np.uint64(1)
    else:
        absent = np.arange(1 << 14, (1 << 14) + 1000, dtype=np.uint64)
    q = np.concatenate([present, absent])
    b, c = ctx.index_lookup(idx, d, q)
    lo, cnt = np_runs(keys, q, shift)
    assert np.array_equal(c.astype(np.int64), cnt)
    hit = cnt > 0
    assert np.array_equal(b[hit].astype(np.int64), lo[hit])
    ctx.index_free(idx)
    d.free()
    ctx.eng.gdl_set_load_factor(ctx.h, 0.35)


def test_index_saturated_run_count(ctx):
    # one prefix with more than 2^24 - 1 tuples exercises the saturated count
    n = (1 << 24) + 5
    keys = np.concatenate([gdl.pack(np.full(n, 5), np.arange(n)),
                           gdl.pack(np.full(3, 9), np.arange(3))])
    d = ctx.to_device(keys)
    idx = ctx.index_build(d, 1)
    b, c = ctx.index_lookup(idx, d, np.array([5, 9, 7], dtype=np.uint64))
    assert c.tolist() == [n, 3, 0] and b[:2].tolist() == [0, n]
    ctx.index_free(idx)
    d.free()


def np_join(inner, outer, colmap, prefix_cols):
    shift = np.uint64(32 if prefix_cols == 1 else 0)
    lo, cnt = np_runs(inner, outer >> shift, int(shift))
    oi = np.repeat(np.arange(outer.size), cnt)
    start = np.repeat(lo, cnt)
    excl = np.repeat(np.cumsum(cnt) - cnt, cnt)
    ii = start + (np.arange(oi.size) - excl)
    ik, ok = inner[ii], outer[oi]
    cols = [ik >> np.uint64(32), ik & M32, ok >> np.uint64(32), ok & M32]
    out = cols[colmap[0]] << np.uint64(32)
    if len(colmap) == 2:
        out = out | cols[colmap[1]]
    return out


@pytest.mark.parametrize("colmap", [(1, 3), (3, 1), (0, 2), (2,), (3, 3)])
@pytest.mark.parametrize("n_outer", [0, 1, 7000, 400_000])
def test_join_matches_numpy(ctx, colmap, n_outer):
    inner = np.unique(rand_keys(50_000, 12, 14))
    outer = rand_keys(n_outer, 12, 14)  # unsorted, with repeats
    got = ctx.join(inner, outer, colmap, 1)
    want = np_join(inner, outer, colmap, 1)
    assert got.size == want.size
    assert np.array_equal(np.sort(got), np.sort(want))


def test_join_skewed_hub(ctx):
    # one join key with 100k matches (the CSPA hub case)
    inner = np.unique(np.concatenate([gdl.pack(np.full(100_000, 3), np.arange(100_000)),
                                      rand_keys(20_000, 10, 10)]))
    outer = gdl.pack(np.array([3, 3, 4, 3]), np.array([1, 2, 3, 4]))
    got = ctx.join(inner, outer, (1, 3), 1)
    want = np_join(inner, outer, (1, 3), 1)
    assert np.array_equal(np.sort(got), np.sort(want))


def test_join_two_column_prefix(ctx):
    inner = np.unique(rand_keys(30_000, 8, 8))
    outer = rand_keys(30_000, 8, 8)
    got = ctx.join(inner, outer, (2, 1), 2)
    want = np_join(inner, outer, (2, 1), 2)
    assert np.array_equal(np.sort(got), np.sort(want))


@pytest.mark.parametrize("op,fn", [(gdl.PRED_NE, np.not_equal), (gdl.PRED_EQ, np.equal),
                                   (gdl.PRED_LT, np.less), (gdl.PRED_GE, np.greater_equal)])
def test_join_with_predicate(ctx, op, fn):
    inner = np.unique(rand_keys(20_000, 8, 6))
    outer = rand_keys(20_000, 8, 6)
    got = ctx.join(inner, outer, (1, 3), 1, (op, 0, 1))
    want = np_join(inner, outer, (1, 3), 1)
    want = want[fn(want >> np.uint64(32), want & M32)]
    assert np.array_equal(np.sort(got), np.sort(want))


def test_project_and_filter(ctx):
    k = rand_keys(123_457, 20, 20)
    assert np.array_equal(ctx.project(k, (1, 0)), ((k & M32) << np.uint64(32)) | (k >> np.uint64(32)))
    assert np.array_equal(ctx.project(k, (0, 0)), ((k >> np.uint64(32)) << np.uint64(32)) | (k >> np.uint64(32)))
    assert np.array_equal(ctx.project(k, (1,)), (k & M32) << np.uint64(32))
    kk = rand_keys(200_000, 4, 4)
    f = ctx.filter(kk, gdl.PRED_NE)
    assert np.array_equal(f, kk[(kk >> np.uint64(32)) != (kk & M32)])


# sorted keys put many lanes of a wave on one bitmap word (and repeats on
# neighbouring lanes): the bitmap insert aggregates them per word
@pytest.mark.parametrize("order", ["random", "sorted"])
@pytest.mark.parametrize("set_kind", ["hash", "bitmap"])
def test_hashset_insert_new_is_set_difference(ctx, set_kind, order):
    hs = ctx.hashset_create(0) if set_kind == "hash" else ctx.hashset_create_bitmap(10)
    seen = set()
    try:
        for batch in range(6):  # the hash table grows and rehashes across batches
            k = rand_keys(150_000 * (batch + 1), 10, 10)  # heavy repeats
            if order == "sorted":
                k = np.sort(k)
            if batch == 2 and set_kind == "hash":
                k = np.concatenate([k, np.array([0xffffffffffffffff] * 3, dtype=np.uint64)])
            new = ctx.hashset_insert_new(hs, k)
            want = set(k.tolist()) - seen
            assert new.size == len(want)
            assert set(new.tolist()) == want
            seen |= want
        assert hs.size == len(seen)
        assert np.array_equal(np.sort(ctx.hashset_extract(hs)), np.array(sorted(seen), np.uint64))
        assert ctx.hashset_check(hs) == 0
    finally:
        ctx.hashset_free(hs)


def test_bitmap_set_refuses_keys_outside_its_domain(ctx):
    hs = ctx.hashset_create_bitmap(8)  # ids 0..255 in both columns
    try:
        inside = gdl.pack([0, 255, 17], [255, 0, 17])
        assert set(ctx.hashset_insert_new(hs, inside).tolist()) == set(inside.tolist())
        assert ctx.hashset_check(hs) == 0
        outside = gdl.pack([256, 3], [1, 1 << 20])
        assert ctx.hashset_insert_new(hs, outside).size == 0  # not inserted ...
        assert ctx.hashset_check(hs) == gdl.GDL_ERR_OVERFLOW  # ... and reported
        assert np.array_equal(np.sort(ctx.hashset_extract(hs)), np.sort(inside))
    finally:
        ctx.hashset_free(hs)
    with pytest.raises(gdl.GdlError):
        ctx.hashset_create_bitmap(19)  # above 8 GiB


def test_allocator_cap_reports_out_of_memory(ctx):
    ctx.set_mem_cap(1 << 20)
    try:
        with pytest.raises(gdl.GdlError) as e:
            ctx.sort(rand_keys(1_000_000))
        assert e.value.status == 2
    finally:
        ctx.set_mem_cap(0)
    # the context still works after the failure
    k = rand_keys(10_000)
    assert np.array_equal(ctx.sort(k), np.sort(k))


def test_allocator_cache_limit_and_best_fit_reuse():
    with gdl.Context(device=0) as c:
        c.release_cached()
        # a freed block is reused for a request up to 2x smaller, not larger
        a = c.alloc(1 << 20)
        a.free()
        assert c.mem_info()["cached"] >= 8 << 20
        b = c.alloc(600_000)  # 4.8 MB <= 8 MB <= 9.6 MB: the cached block
        assert c.mem_info()["cached"] == 0
        b.free()
        # past the limit, freed blocks go back (largest first): 8 + 4 MB
        # exceed 10 MB, the 8 MB block goes, then 2 and 1 MB fit beside 4
        c.set_cache_limit(10 << 20)
        blocks = [c.alloc(n) for n in (1 << 20, 1 << 19, 1 << 18, 1 << 17)]
        for x in blocks:
            x.free()
        info = c.mem_info()
        assert info["in_use"] == 0 and info["cached"] == 7 << 20
        c.set_cache_limit(0)
        # the context still computes after its cache was trimmed
        k = rand_keys(50_000)
        assert np.array_equal(c.sort(k), np.sort(k))


@pytest.mark.parametrize("colmap,pred", [((1, 3), (gdl.PRED_NONE, 0, 1)),
                                         ((3, 1), (gdl.PRED_NE, 0, 1))])
@pytest.mark.parametrize("inner_kind", ["uniform", "hub"])
@pytest.mark.parametrize("set_kind", ["hash", "bitmap"])
def test_join_insert_is_join_then_set_difference(ctx, colmap, pred, inner_kind, set_kind):
    if inner_kind == "hub":  # runs above the wave-cooperative threshold
        inner = np.unique(np.concatenate([gdl.pack(np.full(5000, 7), np.arange(5000)),
                                          rand_keys(20_000, 9, 9)]))
    else:
        inner = np.unique(rand_keys(50_000, 10, 12))
    hs = ctx.hashset_create(0) if set_kind == "hash" else ctx.hashset_create_bitmap(13)
    try:
        seen = set()
        for step in range(3):
            outer = rand_keys(30_000 * (step + 1), 10, 12)
            if inner_kind == "hub":
                outer = np.concatenate([outer, gdl.pack([7, 7], [1, 2])])
            want_all = np_join(inner, outer, colmap, 1)
            if pred[0] == gdl.PRED_NE:
                want_all = want_all[(want_all >> np.uint64(32)) != (want_all & M32)]
            new, raw = ctx.join_insert(hs, inner, outer, colmap, 1, pred)
            assert raw == want_all.size
            want_new = set(want_all.tolist()) - seen
            assert new.size == len(want_new) and set(new.tolist()) == want_new
            seen |= want_new
    finally:
        ctx.hashset_free(hs)


def test_graph_capture_replays_kernel_only_calls(ctx):
    """A captured projection, column mask and segment labelling replayed three
    times, with the outputs overwritten and other eager work in between; every
    replay must equal the eager run."""
    ctx.stats_enable(False)  # kernel timing events cannot be captured
    keys = rand_keys(200_000)
    n = keys.size
    d_in = ctx.to_device(keys)
    d_proj, d_mask = ctx.alloc(n), ctx.alloc(n)
    d_lab = ctx.alloc(n, np.uint32)
    bounds = np.array([0, 1000, 50_000, n], dtype=np.uint64)
    d_b = ctx.to_device(bounds)
    cm = gdl.ColMap(2, (gdl.COL_INNER1, gdl.COL_INNER0))

    def calls():
        ctx._check(ctx.eng.gdl_project_u64(ctx.h, d_in.ptr, n, cm, d_proj.ptr), "project")
        ctx._check(ctx.eng.gdl_mask_columns_u64(ctx.h, d_proj.ptr, n, 1, d_mask.ptr), "mask")
        ctx._check(ctx.eng.gdl_segment_labels_u32(ctx.h, d_b.ptr, 3, 7, d_lab.ptr, n), "labels")

    calls()
    ctx.sync()
    want = (d_proj.numpy(), d_mask.numpy(), d_lab.numpy())
    swapped = (keys << np.uint64(32)) | (keys >> np.uint64(32))
    assert np.array_equal(want[0], swapped)
    assert np.array_equal(want[2], np.repeat(np.array([7, 8, 9], dtype=np.uint32),
                                             np.diff(bounds).astype(np.int64)))
    ctx.graph_begin()
    with pytest.raises(gdl.GdlError):
        ctx.graph_begin()  # one capture at a time
    calls()
    g = ctx.graph_end()
    try:
        junk = np.full(n, 0xDEADBEEF, dtype=np.uint64)
        for rep in range(3):
            for d in (d_proj, d_mask):
                ctx._check(ctx.eng.gdl_memcpy_h2d(ctx.h, d.ptr, junk.ctypes.data, junk.nbytes), "h2d")
            ctx._check(ctx.eng.gdl_memcpy_h2d(ctx.h, d_lab.ptr, junk.ctypes.data, n * 4), "h2d")
            assert np.array_equal(ctx.sort(keys[: 1000 * (rep + 1)]),
                                  np.sort(keys[: 1000 * (rep + 1)]))  # eager work between
            ctx.graph_launch(g)
            ctx.sync()
            got = (d_proj.numpy(), d_mask.numpy(), d_lab.numpy())
            for a, b in zip(got, want):
                assert np.array_equal(a, b), rep
    finally:
        ctx.graph_destroy(g)
        for d in (d_in, d_proj, d_mask, d_lab, d_b):
            d.free()


@pytest.mark.parametrize("bits,n_seg,first", [(16, 1000, 1), (31, 1, 1), (20, 3, 5), (8, 40, 1)])
def test_segment_labels_ride_through_the_sort(ctx, bits, n_seg, first):
    """Distinct keys in log order, labelled by segment, come out sorted with
    each key's label; ids of `bits` bits plus the label's bits fill at most 32."""
    rng = np.random.default_rng(bits * 1000 + n_seg)
    n = min(60_000, (1 << bits) * 4)
    keys = np.unique(rand_keys(n, bits, bits, rng))
    rng.shuffle(keys)
    n = keys.size
    cuts = np.sort(rng.choice(np.arange(1, n), size=n_seg - 1, replace=False)) if n_seg > 1 \
        else np.zeros(0, dtype=np.int64)
    bounds = np.concatenate([[0], cuts, [n]]).astype(np.uint64)
    labels = (first + np.repeat(np.arange(n_seg), np.diff(bounds).astype(np.int64))).astype(np.uint32)
    order = np.argsort(keys, kind="stable")
    d_log, d_b = ctx.to_device(keys), ctx.to_device(bounds)
    d_k, d_l = ctx.alloc(n), ctx.alloc(n, np.uint32)
    try:
        ctx._check(ctx.eng.gdl_segment_labels_sorted(ctx.h, d_log.ptr, n, d_b.ptr, n_seg, first,
                                                      bits, d_k.ptr, d_l.ptr), "labels_sorted")
        assert np.array_equal(d_k.numpy(), keys[order])
        assert np.array_equal(d_l.numpy(), labels[order])
    finally:
        for d in (d_log, d_b, d_k, d_l):
            d.free()


def test_segment_labels_too_wide_for_the_ids_are_refused(ctx):
    d = ctx.to_device(np.array([1, 2], dtype=np.uint64))
    b = ctx.to_device(np.array([0, 1, 2], dtype=np.uint64))
    o, lab = ctx.alloc(2), ctx.alloc(2, np.uint32)
    try:  # 31-bit ids and labels up to 2: 2 label bits make 33
        st = ctx.eng.gdl_segment_labels_sorted(ctx.h, d.ptr, 2, b.ptr, 2, 1, 31, o.ptr, lab.ptr)
        assert st == 4
    finally:
        for x in (d, b, o, lab):
            x.free()
