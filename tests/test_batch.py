"""Batched TC fixpoints (include/gdl_batch.h): every instance of a batch gives
exactly what the TC program gives alone -- the sorted path set, the
per-iteration |delta| and |full|, the iteration count -- whatever ran beside
it, with admission deferred by slots and by the memory budget.

Alone results come from the oracle (the CPU restatement of the reference's
semi-naive loop) and from the GPU program run of the same instance."""
import numpy as np
import pytest

import oracle
from conftest import gdl, read_edges


def small_instances(n, lo=200, hi=3000, seed0=100):
    """(qid, edges, n_nodes) of the batched-TC generator at reduced edge counts."""
    out = []
    for q in range(n):
        seed = seed0 + q
        _, m, nodes = gdl.gen_batched_shape(seed, lo, hi)
        out.append((q, gdl.gen_batched_instance(seed, lo, hi, 100), nodes))
    return out


def delta_full(trace):
    return [(t[0], t[3], t[4]) for t in trace]


# ---------------------------------------------------------------- CPU side

def test_tc_estimate_is_exact_with_every_source():
    for edges, n in [(gdl.gen_er(5, 300, 900), 300), (gdl.gen_tree(2, 2000, 50), 2000),
                     (gdl.pack(np.arange(99), np.arange(1, 100)), 100),
                     (gdl.pack([0, 1, 2], [1, 2, 0]), 3)]:
        want = oracle.tc(edges)["n_out"]
        assert gdl.tc_estimate(edges, n, samples=n) == want


def test_tc_estimate_sampled_is_close_on_er_and_trees():
    e = gdl.gen_er(0, 2000, 5000)
    est = gdl.tc_estimate(e, 2000, samples=64, seed=1)
    assert abs(est - 3_134_855) / 3_134_855 < 0.25
    t = gdl.gen_tree(0, 20000, 1000)
    est = gdl.tc_estimate(t, 20000, samples=64, seed=1)
    want = oracle.tc(t)["n_out"]
    assert abs(est - want) / want < 0.25


# ---------------------------------------------------------------- GPU side

@pytest.mark.gpu
def test_key_histogram_and_mask_columns(ctx):
    rng = np.random.default_rng(5)
    slot = rng.integers(0, 200, 300_000, dtype=np.uint64)
    slot[:70_000] = 17  # long single-bin stretches take the wave-aggregated add
    a = rng.integers(0, 1 << 16, slot.size, dtype=np.uint64)
    c = rng.integers(0, 1 << 16, slot.size, dtype=np.uint64)
    keys = (((slot << np.uint64(16)) | a) << np.uint64(32)) | ((slot << np.uint64(16)) | c)
    h = ctx.key_histogram(keys, 48, 256)
    assert np.array_equal(h, np.bincount(slot.astype(np.int64), minlength=256).astype(np.uint64))
    h = ctx.key_histogram(keys, 48, 100)  # bins past n_bins are not counted
    assert np.array_equal(h, np.bincount(slot.astype(np.int64), minlength=256)[:100])
    assert np.array_equal(ctx.mask_columns(keys, 0xffff), (a << np.uint64(32)) | c)


@pytest.mark.gpu
def test_hashset_extract_returns_the_members(ctx):
    rng = np.random.default_rng(6)
    k = rng.integers(0, 1 << 40, 200_000, dtype=np.uint64)
    k = np.concatenate([k, np.array([0xffffffffffffffff], dtype=np.uint64)])
    hs = ctx.hashset_create(0)
    try:
        ctx.hashset_insert_new(hs, k)
        got = ctx.hashset_extract(hs)
        assert np.array_equal(np.sort(got), np.unique(k))
    finally:
        ctx.hashset_free(hs)


@pytest.mark.gpu
@pytest.mark.parametrize("hub", [False, True])
def test_join_insert_seg_is_per_instance_join_then_difference(ctx, hub):
    rng = np.random.default_rng(7)
    shift = 48
    inner, outer, sets = [], [], []
    for s in range(3):
        pre = np.uint64(s << 16)
        a = rng.integers(0, 400, 6000, dtype=np.uint64) | pre
        b = rng.integers(0, 400, 6000, dtype=np.uint64) | pre
        inner.append((a << np.uint64(32)) | b)
        if hub:  # a run above the load-balancing threshold
            inner.append(((pre | np.uint64(3)) << np.uint64(32)) |
                         (pre | np.arange(5000, dtype=np.uint64)))
        x = rng.integers(0, 400, 4000, dtype=np.uint64) | pre
        y = rng.integers(0, 400, 4000, dtype=np.uint64) | pre
        outer.append((x << np.uint64(32)) | y)
    inner = np.unique(np.concatenate(inner))
    outer = np.concatenate(outer)
    seen = [set(), set(), set()]
    sets = [ctx.hashset_create(1 << 21) for _ in range(3)]
    try:
        for rnd in range(2):
            o = outer[rnd::2]
            new, raw = ctx.join_insert_seg(sets + [None], shift, inner, o, (1, 3))
            # brute force: runs of inner keyed by col0
            ih = inner >> np.uint64(32)
            lo = np.searchsorted(ih, o >> np.uint64(32), "left")
            hi = np.searchsorted(ih, o >> np.uint64(32), "right")
            prod = np.concatenate([((inner[l:h] & np.uint64(0xffffffff)) << np.uint64(32)) |
                                   (v & np.uint64(0xffffffff))
                                   for l, h, v in zip(lo, hi, o)] or
                                  [np.zeros(0, np.uint64)])
            assert raw == prod.size
            for s in range(3):
                mine = prod[(prod >> np.uint64(shift)) == np.uint64(s)]
                want = set(mine.tolist()) - seen[s]
                got = new[(new >> np.uint64(shift)) == np.uint64(s)]
                assert got.size == len(want) and set(got.tolist()) == want
                seen[s] |= want
    finally:
        for hs in sets:
            ctx.hashset_free(hs)


@pytest.mark.gpu
def test_batch_matches_alone_runs_with_slot_reuse(ctx):
    inst = small_instances(20)
    inst.append((20, read_edges("data_7035.txt"), 6105))  # repeated rows: raw-prefix delta
    inst.append((21, gdl.pack([1], [2]), 3))
    with gdl.TcBatch(ctx, max_slots=6) as b:  # 22 instances through 6 slots
        done = b.run(inst)
        info = b.info()
    assert info["admitted"] == info["finished"] == len(inst)
    for qid, edges, n in inst:
        r = done[qid]
        want = oracle.tc(edges)
        assert r["n_out"] == want["n_out"], qid
        assert np.array_equal(r["keys"], want["out"]), qid
        assert r["iterations"] == want["iterations"], qid
        assert delta_full(r["trace"]) == delta_full(want["trace"]), qid
        assert r["derived"] == want["derived"], qid
        assert r["checksum"] == gdl.tuple_checksum(want["out"]), qid


@pytest.mark.gpu
@pytest.mark.parametrize("solo", [0, 20_000])
def test_solo_lane_matches_the_oracle(ctx, solo):
    # solo 0: every instance runs alone at admission; 20k: the larger ones do
    inst = small_instances(14)
    inst.append((14, read_edges("data_7035.txt"), 6105))  # repeated rows: raw-prefix delta
    inst.append((15, gdl.pack([1], [2]), 3))
    inst.append((16, gdl.pack([0, 1, 2], [1, 2, 0]), 3))  # a cycle: body 0 adds tuples
    with gdl.TcBatch(ctx, max_slots=5, solo_tuples=solo) as b:
        done = b.run(inst)
        info = b.info()
    assert info["finished"] == len(inst)
    slots = [done[q]["slot"] for q, _, _ in inst]
    assert (min(slots) == max(slots) == -1) if solo == 0 else (-1 in slots and max(slots) >= 0)
    for qid, edges, n in inst:
        r = done[qid]
        want = oracle.tc(edges)
        assert r["n_out"] == want["n_out"], qid
        assert np.array_equal(r["keys"], want["out"]), qid
        assert r["iterations"] == want["iterations"], qid
        assert delta_full(r["trace"]) == delta_full(want["trace"]), qid
        assert r["derived"] == want["derived"], qid
        assert r["checksum"] == gdl.tuple_checksum(want["out"]), qid


@pytest.mark.gpu
def test_batch_alone_vs_mix_on_the_gpu(ctx):
    inst = small_instances(12, 1000, 20000, seed0=0)
    with gdl.TcBatch(ctx, keep_keys=False) as b:
        mix = b.run(inst)
    for qid, edges, n in inst:
        alone = ctx.run_tc(edges)
        r = mix[qid]
        assert r["keys"] is None
        assert r["n_out"] == alone["n_out"] and r["iterations"] == alone["iterations"]
        assert r["checksum"] == gdl.tuple_checksum(alone["out"])
        assert delta_full(r["trace"]) == delta_full(alone["trace"])
        # the same instance alone in a batch of one
        with gdl.TcBatch(ctx, max_slots=1, keep_keys=False) as one:
            solo = one.run([(qid, edges, n)])[qid]
        assert (solo["n_out"], solo["checksum"], solo["iterations"], solo["trace"]) == \
            (r["n_out"], r["checksum"], r["iterations"], r["trace"])


@pytest.mark.gpu
def test_batch_budget_defers_admission(ctx):
    inst = small_instances(6)
    with gdl.TcBatch(ctx, mem_budget=1) as b:  # room for one instance at a time
        assert b.admit(*inst[0])
        assert not b.admit(*inst[1])  # busy: the budget is taken
        pending, done = list(inst[1:]), {}
        while len(done) < len(inst):
            while pending and b.admit(*pending[0]):
                pending.pop(0)
            assert b.step() <= 1
            done.update({r["qid"]: r for r in b.poll()})
        assert b.info()["finished"] == len(inst)
    for qid, edges, n in inst:
        want = oracle.tc(edges)
        assert np.array_equal(done[qid]["keys"], want["out"])
        assert done[qid]["iterations"] == want["iterations"]


@pytest.mark.gpu
def test_batch_rejects_ids_outside_the_slot(ctx):
    with gdl.TcBatch(ctx, max_slots=4, node_bits=8) as b:
        with pytest.raises(gdl.GdlError) as e:
            b.admit(0, gdl.pack([1], [300]), 256)
        assert e.value.status == 1
        with pytest.raises(gdl.GdlError):
            b.admit(1, gdl.pack([1], [2]), 257)  # more nodes than 2^node_bits


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", [dict(radix_bits=6, join_block=64, blocks_per_cu=4),
                                    dict(radix_bits=7, join_block=512, blocks_per_cu=16,
                                         load_factor=0.8),
                                    dict(radix_bits=5, join_block=128, load_factor=0.7)])
def test_tuning_keeps_results(tuning):
    # a context of its own: tuning is per context
    with gdl.Context(device=0) as c:
        c.set_tuning(**tuning)
        got = c.tuning()
        for k, v in tuning.items():
            assert abs(got[k] - v) < 1e-6
        rng = np.random.default_rng(3)
        k = rng.integers(0, 1 << 62, 300_000, dtype=np.uint64)
        assert np.array_equal(c.sort(k, 32, 32), np.sort(k))
        k = (rng.integers(0, 1 << 13, 5000, dtype=np.uint64) << np.uint64(32)) | \
            rng.integers(0, 1 << 11, 5000, dtype=np.uint64)
        assert np.array_equal(c.sort_unique(k, 13, 11), np.unique(k))
        for edges in (gdl.gen_er(3, 500, 1500), gdl.gen_tree(11, 5000, 100)):
            want = oracle.tc(edges)
            r = c.run_tc(edges)
            assert np.array_equal(r["out"], want["out"]) and r["iterations"] == want["iterations"]
            assert r["checksum"] == gdl.tuple_checksum(want["out"])
        with gdl.TcBatch(c, max_slots=4) as b:
            inst = small_instances(6)
            done = b.run(inst)
        for qid, edges, _ in inst:
            assert np.array_equal(done[qid]["keys"], oracle.tc(edges)["out"])


@pytest.mark.gpu
def test_autotune_picks_from_the_grid_and_caches(tmp_path):
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "gdl_tune", os.path.join(os.path.dirname(gdl.__file__), "tune.py"))
    tune = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(tune)
    inst = small_instances(4)
    with gdl.Context(device=0) as c:
        def work():
            with gdl.TcBatch(c, keep_keys=False) as b:
                b.run(inst)
        p = str(tmp_path / "tuning.json")
        first = tune.autotune(c, work, key="t", reps=1, path=p)
        assert not first["cached"] and first["trials"] >= 9
        for name, values in tune.SPACE.items():
            assert first["tuning"][name] in values
        again = tune.autotune(c, work, key="t", path=p)
        assert again["cached"] and again["tuning"] == first["tuning"]
        assert c.tuning()["join_block"] == first["tuning"]["join_block"]
