"""Bodies on bit matrices (gdl_fixpoint_dense, csrc/dense.hip) and the
key-ordered extraction of sets.

The rule shapes are TC's (out(x, z) :- r1(y, x), delta(y, z); test/tc.cu)
and SG's (out(x, y) :- r1(a, x), delta(a, b), r2(b, y) through a
deduplicated tmp; test/sg.cu).  `tc_bodies` / `sg_bodies` below restate one
semi-naive body per loop with Python sets; they are pinned against the
oracle's TC and SG traces (new tuples and |FULL| per iteration) on the CPU.
On the GPU, the dense bodies must give the same per-body new-tuple and
produced counts, FULL and log; whole programs must give the same results and
traces (produced counts included) with dense bodies off, chosen by the cost
model, and forced."""
from collections import defaultdict

import numpy as np
import pytest

import oracle
from conftest import gdl, read_edges

M32 = np.uint64(0xffffffff)


def pairs(keys):
    k = np.asarray(keys, dtype=np.uint64)
    return list(zip((k >> np.uint64(32)).tolist(), (k & M32).tolist()))


def key(a, b):
    return (a << 32) | b


def tc_bodies(r1, full, delta, max_bodies=1 << 30):
    """Bodies of out(x, z) :- r1(y, x), delta(y, z): per body (new, produced);
    the produced count runs over the delta as given (repeats included)."""
    by_y = defaultdict(list)
    for y, x in pairs(r1):
        by_y[y].append(x)
    full = set(int(k) for k in full)
    d = pairs(delta)
    out = []
    while len(out) < max_bodies:
        raw, new = 0, set()
        for y, z in d:
            for x in by_y.get(y, ()):
                raw += 1
                k = key(x, z)
                if k not in full:
                    new.add(k)
        full |= new
        out.append((len(new), raw))
        if not new:
            break
        d = pairs(sorted(new))
    return out, full


def sg_bodies(r1, r2, full, delta, max_bodies=1 << 30):
    """Bodies of out(x, y) :- r1(a, x), delta(a, b), r2(b, y) through the
    deduplicated tmp(b, x); produced = the second join's over that tmp."""
    by_a, by_b = defaultdict(list), defaultdict(list)
    for a, x in pairs(r1):
        by_a[a].append(x)
    for b, y in pairs(r2):
        by_b[b].append(y)
    full = set(int(k) for k in full)
    d = pairs(delta)
    out = []
    while len(out) < max_bodies:
        tmp = {(b, x) for a, b in d for x in by_a.get(a, ())}
        raw = sum(len(by_b.get(b, ())) for b, _ in tmp)
        new = {key(x, y) for b, x in tmp for y in by_b.get(b, ())} - full
        full |= new
        out.append((len(new), raw))
        if not new:
            break
        d = pairs(sorted(new))
    return out, full


def reverse(keys):
    k = np.asarray(keys, dtype=np.uint64)
    return (k << np.uint64(32)) | (k >> np.uint64(32))


GRAPHS = {
    "er_300": lambda: gdl.gen_er(3, 300, 900),
    "er_1000": lambda: gdl.gen_er(4, 1000, 2600),
    "tree_2k": lambda: gdl.gen_tree(5, 2000, 40),
    "rmat_10": lambda: gdl.gen_rmat(6, 10, 2500),
    "hpc_talk": lambda: read_edges("hpc_talk.txt"),
}


@pytest.mark.parametrize("name", ["er_300", "tree_2k", "hpc_talk"])
def test_tc_bodies_restate_the_oracle(name):
    e = np.unique(GRAPHS[name]())
    want = oracle.tc(e)
    bodies, full = tc_bodies(reverse(e), e, e)
    assert [b[0] for b in bodies] == [t[3] for t in want["trace"]]
    assert sorted(full) == want["out"].tolist()
    # the full column of the trace
    f = e.size
    for (n, _), t in zip(bodies, want["trace"]):
        f += n
        assert f == t[4]


@pytest.mark.parametrize("name", ["er_300", "rmat_10", "hpc_talk"])
def test_sg_bodies_restate_the_oracle(name):
    e = np.unique(GRAPHS[name]())
    want = oracle.sg(e)
    init = np.unique(np.array(
        [key(x, y) for p, xs in _by_first(e).items() for x in xs for y in xs if x != y],
        dtype=np.uint64))
    assert init.size == want["trace_init"][0][3]
    bodies, full = sg_bodies(e, e, init, init)
    assert [b[0] for b in bodies] == [t[3] for t in want["trace"]]
    assert sorted(full) == want["out"].tolist()


def _by_first(e):
    d = defaultdict(list)
    for a, b in pairs(e):
        d[a].append(b)
    return d


def test_produced_counts_include_delta_repeats():
    # r1(y, x): 1 -> {5, 6}; delta (1, 9) twice: produced 4, new 2
    bodies, _ = tc_bodies([key(1, 5), key(1, 6)], [key(1, 9)], [key(1, 9), key(1, 9)], 1)
    assert bodies == [(2, 4)]


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("kb", [2, 5, 7, 12, 15])
def test_bitmap_extract_sorted(ctx, kb):
    rng = np.random.default_rng(kb)
    n = min(50_000, 1 << (2 * kb))
    k = (rng.integers(0, 1 << kb, n, dtype=np.uint64) << np.uint64(32)) | \
        rng.integers(0, 1 << kb, n, dtype=np.uint64)
    hs = ctx.hashset_create_bitmap(kb)
    try:
        ctx.hashset_insert_new(hs, k)
        assert np.array_equal(ctx.hashset_extract_sorted(hs), np.unique(k))
    finally:
        ctx.hashset_free(hs)


@pytest.mark.gpu
def test_hash_set_extract_sorted(ctx):
    rng = np.random.default_rng(1)
    k = rng.integers(0, 1 << 63, 100_000, dtype=np.uint64)
    hs = ctx.hashset_create(1000)
    try:
        ctx.hashset_insert_new(hs, k)
        assert np.array_equal(ctx.hashset_extract_sorted(hs), np.unique(k))
    finally:
        ctx.hashset_free(hs)


def _check_run(got, bodies, full0, full_set):
    assert list(zip(got["new"], got["raw"])) == bodies
    assert got["members"].tolist() == sorted(full_set)
    assert got["size"] == len(full_set)
    log = got["log"]
    assert log.size == len(full_set)
    assert np.array_equal(np.sort(log[:full0.size]), np.sort(full0))
    at = full0.size
    prev = set(full0.tolist())
    for n, _ in bodies:  # each body's new tuples, as one segment of the log
        seg = set(log[at:at + n].tolist())
        assert len(seg) == n and not (seg & prev)
        prev |= seg
        at += n


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_dense_tc_bodies(ctx, name):
    e = np.unique(GRAPHS[name]())
    kb = max(6, int(max(int((e >> np.uint64(32)).max()), int((e & M32).max())) + 1).bit_length())
    r1 = reverse(e)
    # a delta with repeats: every edge, then the first tenth again
    delta = np.concatenate([e, e[: e.size // 10]])
    want, fs = tc_bodies(r1, e, delta)
    ctx.set_dense_mode(2)
    try:
        got = ctx.fixpoint_dense(gdl.DENSE_TC, kb, e, delta, r1)
    finally:
        ctx.set_dense_mode(1)
    _check_run(got, want, e, fs)


@pytest.mark.gpu
def test_dense_tc_bodies_delta_rows_without_preds(ctx):
    # delta rows that no r1 tuple points at: the first body reads them, no
    # later body writes them, so the third body (which reads the first
    # body's buffers again) must see them empty
    e = np.unique(GRAPHS["er_1000"]())
    src = set((e >> np.uint64(32)).tolist())
    lonely = np.array([v for v in range(1024) if v not in src][:40], dtype=np.uint64)
    extra = (lonely << np.uint64(32)) | np.arange(lonely.size, dtype=np.uint64)
    delta = np.unique(np.concatenate([e, extra]))
    r1 = reverse(e)
    want, fs = tc_bodies(r1, delta, delta, max_bodies=4)
    ctx.set_dense_mode(2)
    try:
        got = ctx.fixpoint_dense(gdl.DENSE_TC, 10, delta, delta, r1, max_bodies=4)
    finally:
        ctx.set_dense_mode(1)
    _check_run(got, want, delta, fs)


# dense mode 2: the cost model picks how each SG body expands U = D R2;
# 3: always by columns (two bit-matrix transposes around a row OR over r2's
# pred lists, for domains of 2^9 ids or more); 4: always bit by bit
@pytest.mark.gpu
@pytest.mark.parametrize("mode", [2, 3, 4])
@pytest.mark.parametrize("name", ["er_300", "er_1000", "rmat_10", "hpc_talk"])
def test_dense_sg_bodies(ctx, name, mode):
    e = np.unique(GRAPHS[name]())
    kb = max(6, int(max(int((e >> np.uint64(32)).max()), int((e & M32).max())) + 1).bit_length())
    init = np.unique(np.array(
        [key(x, y) for p, xs in _by_first(e).items() for x in xs for y in xs if x != y],
        dtype=np.uint64))
    want, fs = sg_bodies(e, e, init, init)
    ctx.set_dense_mode(mode)
    try:
        got = ctx.fixpoint_dense(gdl.DENSE_SG, kb, init, init, e, e)
    finally:
        ctx.set_dense_mode(1)
    _check_run(got, want, init, fs)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["tc", "sg"])
def test_dense_log_last_logs_the_last_body_only(ctx, shape):
    # GDL_DENSE_LOG_LAST: the same bodies and set; the log is FULL's start
    # plus the last body's new tuples (the next delta)
    e = np.unique(GRAPHS["er_1000"]())
    if shape == "tc":
        full0 = e
        run = lambda k: tc_bodies(reverse(e), e, e, max_bodies=k)  # noqa: E731
        args = (gdl.DENSE_TC, 10, e, e, reverse(e))
    else:
        full0 = np.unique(np.array(
            [key(x, y) for p, xs in _by_first(e).items() for x in xs for y in xs if x != y],
            dtype=np.uint64))
        run = lambda k: sg_bodies(e, e, full0, full0, max_bodies=k)  # noqa: E731
        args = (gdl.DENSE_SG, 10, full0, full0, e, e)
    want, fs = run(5)
    _, fs_before_last = run(len(want) - 1)
    ctx.set_dense_mode(2)
    try:
        got = ctx.fixpoint_dense(*args, max_bodies=5, flags=gdl.DENSE_LOG_LAST)
    finally:
        ctx.set_dense_mode(1)
    assert list(zip(got["new"], got["raw"])) == want
    assert got["members"].tolist() == sorted(fs)
    log = got["log"]
    last = want[-1][0]
    assert log.size == full0.size + last
    assert np.array_equal(np.sort(log[:full0.size]), np.sort(full0))
    tail = log[full0.size:].tolist()
    assert len(set(tail)) == len(tail) and set(tail) == fs - fs_before_last


@pytest.mark.gpu
@pytest.mark.parametrize("same", [False, True])
@pytest.mark.parametrize("mode", [3, 4])
@pytest.mark.parametrize("kb", [9, 11, 13])
def test_dense_sg_bodies_two_relations(ctx, kb, mode, same):
    # a delta that is not symmetric, with r1 != r2 or with one relation for
    # both (whose symmetric shortcut must then be declined): the column
    # expansion's transposes must not lean on SG's symmetry; r2 has hub
    # columns (a row with more than 32 preds is OR-ed in slices)
    rng = np.random.default_rng(kb)
    n = 1 << kb
    r1 = np.unique(gdl.pack(rng.integers(0, n, 2 * n), rng.integers(0, n, 2 * n)))
    hub = rng.integers(0, n, 8)
    r2 = np.unique(np.concatenate([
        gdl.pack(rng.integers(0, n, 2 * n), rng.integers(0, n, 2 * n)),
        gdl.pack(rng.integers(0, n, 400), np.repeat(hub, 50)),
        gdl.pack(np.repeat(hub[:2], 70), rng.integers(0, n, 140))]))
    if same:
        r1 = r2
    delta = np.unique(gdl.pack(rng.integers(0, n, 2 * n), rng.integers(0, n, 2 * n)))
    want, fs = sg_bodies(r1, r2, delta, delta, max_bodies=3)
    ctx.set_dense_mode(mode)
    try:
        got = ctx.fixpoint_dense(gdl.DENSE_SG, kb, delta, delta, r1, r2, max_bodies=3)
    finally:
        ctx.set_dense_mode(1)
    _check_run(got, want, delta, fs)


@pytest.mark.gpu
def test_dense_bodies_stop_at_max_bodies_and_resume(ctx):
    e = np.unique(GRAPHS["er_1000"]())
    kb = 10
    r1 = reverse(e)
    want, fs = tc_bodies(r1, e, e)
    ctx.set_dense_mode(2)
    try:
        a = ctx.fixpoint_dense(gdl.DENSE_TC, kb, e, e, r1, max_bodies=3)
        assert list(zip(a["new"], a["raw"])) == want[:3]
        # continue from the log's last segment
        last = a["new"][-1]
        b = ctx.fixpoint_dense(gdl.DENSE_TC, kb, a["log"], a["log"][a["log"].size - last:], r1)
    finally:
        ctx.set_dense_mode(1)
    assert list(zip(b["new"], b["raw"])) == want[3:]
    assert b["members"].tolist() == sorted(fs)


@pytest.mark.gpu
def test_dense_mode_zero_runs_no_body(ctx):
    e = np.unique(GRAPHS["er_300"]())
    ctx.set_dense_mode(0)
    try:
        got = ctx.fixpoint_dense(gdl.DENSE_TC, 9, e, e, reverse(e))
    finally:
        ctx.set_dense_mode(1)
    assert got["new"] == [] and got["log"].size == e.size


@pytest.mark.gpu
def test_dense_declines_ids_beyond_the_domain(ctx):
    e = np.unique(GRAPHS["er_300"]())
    r1 = np.concatenate([reverse(e), np.array([key(5, 600)], dtype=np.uint64)])
    ctx.set_dense_mode(2)
    try:
        got = ctx.fixpoint_dense(gdl.DENSE_TC, 9, e, e, r1)  # 600 >= 2^9
    finally:
        ctx.set_dense_mode(1)
    assert got["new"] == []


PROGRAMS = {
    "tc_er_2000": ("tc", lambda: gdl.gen_er(0, 2000, 5000)),
    "tc_er_dups": ("tc", lambda: read_edges("data_7035.txt")),
    "tc_tree": ("tc", lambda: gdl.gen_tree(7, 20000, 300)),
    "sg_rmat_12": ("sg", lambda: gdl.gen_rmat(1, 12, 6000)),
    "sg_er_500": ("sg", lambda: gdl.gen_er(2, 500, 1100)),
    # joins of many tuples per new one over a small id domain run on bit
    # matrices (modes 2-4 force them on every join into a bitmap set)
    "cspa_zipf_150": ("cspa", lambda: gdl.gen_cspa(5, 150, 1.1, 200, 100)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(PROGRAMS))
def test_programs_same_with_dense_bodies_off_auto_forced(ctx, name):
    kind, make = PROGRAMS[name]
    e = make()
    if kind == "cspa":
        run = lambda inp: ctx.run_cspa(*inp, out_cap=2_000_000, extra_cap=2_000_000)  # noqa: E731
    else:
        run = ctx.run_tc if kind == "tc" else ctx.run_sg
    res = []
    for mode in (0, 1, 2, 3, 4):
        ctx.set_dense_mode(mode)
        try:
            res.append(run(e))
        finally:
            ctx.set_dense_mode(1)
    for r in res[1:]:
        assert np.array_equal(r["out"], res[0]["out"])
        assert r["iterations"] == res[0]["iterations"]
        assert r["trace"] == res[0]["trace"]  # produced counts included
        assert r["trace_init"] == res[0]["trace_init"]
    want = {"tc": oracle.tc, "sg": oracle.sg}[kind](e) if kind != "cspa" else oracle.cspa(*e)
    assert np.array_equal(res[2]["out"], want["out"])
    if kind == "cspa":
        for r in res:
            assert r["sizes"] == want["sizes"]
            assert all(np.array_equal(g, w) for g, w in zip(r["extra"], want["extra"]))
    assert [(t[0], t[3], t[4]) for t in res[2]["trace"]] == \
        [(t[0], t[3], t[4]) for t in want["trace"]]
