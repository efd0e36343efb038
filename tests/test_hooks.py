"""Device hooks that are not column selections (gdlog/hook_exec.h): the
engine calls them once per produced tuple, as the reference does
(relation.cu:348-358 for a join's generator and predicate, relation.cu:390-403
for a copy hook).  gdl_prog_tc_hooks runs the reference's TC encoding
(test/tc.cu:80-149) with such hooks; the oracle runs the same LIEs with the
same hooks written in Python (the oracle's LIE takes C function pointers),
and the results must agree bit for bit: path sets, per-iteration traces,
iteration counts."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import gdl, read_edges

u64p = ctypes.POINTER(ctypes.c_uint64)
GEN = ctypes.CFUNCTYPE(None, u64p, u64p, u64p)
COPY = ctypes.CFUNCTYPE(None, u64p, u64p)
PRED = ctypes.CFUNCTYPE(ctypes.c_int, u64p)
DELTA, FULL = 0, 1


@GEN
def tc_gen(inner, outer, out):  # test/tc.cu:61-65
    out[0] = inner[1]
    out[1] = outer[1]


@PRED
def not_into_42(t):
    return int(t[1] != 42)


@GEN
def shift_gen(inner, outer, out):
    out[0] = inner[1]
    out[1] = (outer[1] + 1) % 1000


@COPY
def scale_copy(src, dst):
    dst[0] = src[0]
    dst[1] = (src[1] * 7) % 1000


def _lib():
    L = oracle.lib()
    vp = ctypes.c_void_p
    L.orc_rel_new.restype = vp
    L.orc_rel_new.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_rel_load.argtypes = [vp, u64p, ctypes.c_size_t]
    L.orc_rel_size.restype = ctypes.c_size_t
    L.orc_rel_size.argtypes = [vp, ctypes.c_int]
    L.orc_rel_data.restype = u64p
    L.orc_rel_data.argtypes = [vp, ctypes.c_int]
    L.orc_rel_free.argtypes = [vp]
    L.orc_lie_new.restype = vp
    L.orc_lie_free.argtypes = [vp]
    L.orc_lie_add_relation.argtypes = [vp, vp, ctypes.c_int]
    L.orc_lie_add_join.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, GEN,
                                   ctypes.c_void_p]
    L.orc_lie_add_copy.argtypes = [vp, vp, ctypes.c_int, vp, COPY]
    L.orc_lie_set_reload_full.argtypes = [vp, ctypes.c_int]
    L.orc_lie_run.restype = ctypes.c_int
    L.orc_lie_run.argtypes = [vp]
    L.orc_lie_trace.restype = ctypes.c_size_t
    L.orc_lie_trace.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(oracle.TraceRec))]
    return L


def _load(L, name, keys):
    r = L.orc_rel_new(name.encode(), 2, 1, 0)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    L.orc_rel_load(r, k.ctypes.data_as(u64p), k.size)
    return r


def _trace(L, lie):
    p = ctypes.POINTER(oracle.TraceRec)()
    n = L.orc_lie_trace(lie, ctypes.byref(p))
    return [(p[i].iteration, p[i].rel, p[i].newt, p[i].delta, p[i].full) for i in range(n)]


def oracle_tc_hooks(edges, variant):
    """The LIEs of gdl_prog_tc_hooks on the oracle."""
    L = _lib()
    rev = (edges << np.uint64(32)) | (edges >> np.uint64(32))
    edge = _load(L, "edge_2__2_1", rev)
    fwd = path = None
    lies = []
    try:
        trace_init, it_init = [], 0
        if variant == 2:
            fwd = _load(L, "edge_2__1_2", edges)
            path = _load(L, "path_2__1_2", np.zeros(0, np.uint64))
            init = L.orc_lie_new()
            lies.append(init)
            L.orc_lie_add_relation(init, fwd, 1)
            L.orc_lie_add_relation(init, path, 0)
            L.orc_lie_add_copy(init, fwd, FULL, path, scale_copy)
            it_init = L.orc_lie_run(init)
            trace_init = _trace(L, init)
        else:
            path = _load(L, "path_2__1_2", edges)
        lie = L.orc_lie_new()
        lies.append(lie)
        L.orc_lie_add_relation(lie, edge, 1)
        L.orc_lie_add_relation(lie, path, 0)
        gen = shift_gen if variant == 1 else tc_gen
        pred = ctypes.cast(not_into_42, ctypes.c_void_p) if variant == 0 else None
        L.orc_lie_add_join(lie, edge, FULL, path, DELTA, path, gen, pred)
        L.orc_lie_set_reload_full(lie, 0)  # tc_scc.reload_full_flag = false
        it = L.orc_lie_run(lie)
        n = L.orc_rel_size(path, FULL)
        out = np.ctypeslib.as_array(L.orc_rel_data(path, FULL), shape=(n,)).copy() if n else \
            np.zeros(0, np.uint64)
        return dict(out=out, iterations=it, trace=_trace(L, lie), init_iterations=it_init,
                    trace_init=trace_init)
    finally:
        for lie in lies:
            L.orc_lie_free(lie)
        for r in (edge, fwd, path):
            if r:
                L.orc_rel_free(r)


def test_oracle_hooks_do_what_they_say():
    # variant 0 on a graph with edges into 42: no derived path ends at 42
    e = gdl.pack([1, 2, 3, 42, 42], [2, 3, 42, 5, 6])
    r = oracle_tc_hooks(e, 0)
    derived = set(r["out"].tolist()) - set(e.tolist())
    assert derived and all((k & 0xffffffff) != 42 for k in derived)
    # the full closure minus the paths that end at 42 and were not edges
    tc = set(oracle.tc(e)["out"].tolist())
    assert set(r["out"].tolist()) == {k for k in tc if (k & 0xffffffff) != 42} | set(e.tolist())
    # variant 1 on a chain: (a, b) edge, path(b, c) gives (a, c + 1 mod 1000)
    e = gdl.pack([0, 1], [1, 2])
    r = oracle_tc_hooks(e, 1)
    assert set(r["out"].tolist()) == set(e.tolist()) | {(0 << 32) | 3}
    # variant 2: the copy seeds (a, 7b mod 1000)
    e = gdl.pack([0, 5], [5, 200])
    r = oracle_tc_hooks(e, 2)
    seed = {(0 << 32) | 35, (5 << 32) | 400}
    assert set(r["out"].tolist()) >= seed


GRAPHS = {
    "er_400": lambda: gdl.gen_er(4, 400, 1200) ,
    "tree_3k": lambda: gdl.gen_tree(5, 3000, 50),
    "hub_into_42": lambda: np.unique(np.concatenate(
        [gdl.gen_er(6, 300, 700), gdl.pack(np.arange(100), np.full(100, 42)),
         gdl.pack(np.full(50, 42), np.arange(200, 250))])),
    "data_7035_dups": lambda: read_edges("data_7035.txt"),
}


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_device_hooks_match_the_oracle(ctx, name, variant):
    e = GRAPHS[name]()
    if variant > 0:
        e = e[(e & np.uint64(0xffffffff)) < np.uint64(1000)]
    got = ctx.run_tc_hooks(e, variant, out_cap=5_000_000)
    want = oracle_tc_hooks(e, variant)
    assert got["n_out"] == want["out"].size
    assert np.array_equal(got["out"], want["out"])
    assert got["iterations"] == want["iterations"]
    assert [(t[0], t[3], t[4]) for t in got["trace"]] == \
        [(t[0], t[3], t[4]) for t in want["trace"]]
    if variant == 2:
        assert got["init_iterations"] == want["init_iterations"]
        assert [(t[0], t[1], t[3], t[4]) for t in got["trace_init"]] == \
            [(t[0], t[1], t[3], t[4]) for t in want["trace_init"]]
