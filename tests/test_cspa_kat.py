"""Known-answer test of the CSPA program's per-iteration trace, derived by
hand from the reference's operator semantics (not from the oracle):

  * LIE::fixpoint_loop (src/lie.cu:26-216): the initial delta is FULL; RA ops
    run in add_ra order against the FULL of the iteration's start and the
    previous iteration's delta; tmp newts are dropped after the ops
    (lie.cu:85-87); an update relation with an empty newt gets an empty delta
    (lie.cu:114-121); otherwise delta = newt \\ FULL, merged into FULL
    (lie.cu:131-138, relation.cu:405-521); the counter counts bodies,
    including the last empty one (lie.cu:208-215).
  * RelationalJoin (src/join.cu:12-270): an empty outer returns; an empty inner
    zeroes the outer's count; the inner is matched on the outer's first
    column; a tmp_flag output's newt is the raw concatenation of the join
    outputs (join.cu:141-151, 226-254), any other newt is sort+unique'd.
  * RelationalCopy (src/copy.cu:12-95): a FULL source is copied once per LIE
    (lie.cu:69-77); an empty source zeroes the destination newt.
  * RelationalACopy (src/acopy.cu:12-50): the destination newt is overwritten
    with the copied source NEWT of the same iteration (before its set
    difference), or zeroed when that newt is empty.
  * The CSPA encoding: test/cspa.cu:100-379 (14 analysis ops, op indices
    below in add_ra order; tmp_rel_def is loaded without tmp_flag, tmp_rel_ma1
    and tmp_rel_ma2 with it).

Input: Assign file rows (1,2), (5,3); Dereference file rows (1,3), (2,4).
assign_2__2_1 holds the reversed rows {(2,1), (3,5)}.

Init LIE (cspa.cu:192-214), relations VF12, VF21, MA12, MA21:
  it0  the three FULL copies give VF12 newt {(2,2),(3,3)} + {(1,1),(5,5)} +
       {(1,2),(5,3)} = 6; MA12 newt {(2,2),(3,3),(1,1),(5,5)} = 4; the DELTA
       copies see empty deltas (initial delta = empty FULL) -> VF21, MA21 0.
  it1  the FULL copies do not run again, so VF12 / MA12 newts are empty;
       VF21 / MA21 get the reversed deltas (6 and 4).
  it2  every newt is empty -> fixpoint; 3 iterations.
VF0 = {(1,1),(1,2),(2,2),(3,3),(5,3),(5,5)}, MA0 = {(1,1),(2,2),(3,3),(5,5)}.

Analysis LIE, relations VF12, VF21, MA12, MA21, VA (delta = FULL at start):
  it0  op0/op1 (VF o VF) reproduce VF0 (raw 8 each); op2/op3 give VA newt
       {(1,1),(1,2),(2,1),(2,2),(3,3),(3,5),(5,3),(5,5)} (raw 10 each); op4
       (Assign, MA delta) gives (1,2),(5,3) (raw 2); op5/op6 see an empty VA
       delta / tmp_rel_def newt; op7 and op8 each put the same 6 tuples into
       tmp_rel_ma1, whose raw newt is then 12, so op9 joins every tuple twice
       (raw 20); op10 gives 6 into tmp_rel_ma2, op11 raw 10; op12 copies VF12's
       whole newt (6 tuples, all already in FULL) into VF21's newt; op13 sees
       MA12's empty newt and zeroes MA21's.
       VF12 (6,0,6)  VF21 (6,0,6)  MA12 (0,0,4)  MA21 (0,0,4)  VA (8,8,8)
  it1  only VA has a delta: op5 gives tmp_rel_def {(1,3),(2,3),(1,4),(2,4)}
       (raw 4), op6 gives MA12 newt {(3,3),(4,3),(3,4),(4,4)} (raw 4), op13
       mirrors it into MA21.
       VF12 (0,0,6)  VF21 (0,0,6)  MA12 (4,3,7)  MA21 (4,3,7)  VA (0,0,8)
  it2  op4: Assign (3,5) with MA delta (3,4) gives VF12 (5,4) (raw 1); op8:
       VF (3,3) with MA delta (3,4) gives tmp_rel_ma1 (4,3), whose op9 join
       finds no VF row starting at 4 (raw 0); op12 mirrors (4,5).
       VF12 (1,1,7)  VF21 (1,1,7)  MA12 (0,0,7)  MA21 (0,0,7)  VA (0,0,8)
  it3  op0: no VF row starts at 4 (raw 0); op1: VF21 (5,5) with delta (5,4)
       gives (5,4) again (raw 1); op2/op3 give (3,4),(4,4),(5,4) and
       (4,3),(4,4),(4,5) (raw 3 each); op7 -> tmp_rel_ma1 (5,4), op9 raw 3;
       op10 -> tmp_rel_ma2 (5,4), op11 raw 3; op12 mirrors (4,5).
       VF12 (1,0,7)  VF21 (1,0,7)  MA12 (0,0,7)  MA21 (0,0,7)  VA (5,5,13)
  it4  op5 joins the VA delta (sources 3, 4, 5) with Dereference (sources 1,
       2): nothing; every newt is empty -> fixpoint; 5 iterations.
Final: |VF| 7, |MA| 7, |VA| 13 (also the naive cspa.dl fixpoint).
"""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import gdl

ASSIGN = gdl.pack([1, 5], [2, 3])
DEREF = gdl.pack([1, 2], [3, 4])

# (iteration, relation, |newt|, |delta|, |full|); relations in add_relations order
INIT_TRACE = [
    (0, 0, 6, 6, 6), (0, 1, 0, 0, 0), (0, 2, 4, 4, 4), (0, 3, 0, 0, 0),
    (1, 0, 0, 0, 6), (1, 1, 6, 6, 6), (1, 2, 0, 0, 4), (1, 3, 4, 4, 4),
    (2, 0, 0, 0, 6), (2, 1, 0, 0, 6), (2, 2, 0, 0, 4), (2, 3, 0, 0, 4),
]
VF12, VF21, MA12, MA21, VA = range(5)
ANALYSIS_TRACE = [
    (0, VF12, 6, 0, 6), (0, VF21, 6, 0, 6), (0, MA12, 0, 0, 4), (0, MA21, 0, 0, 4),
    (0, VA, 8, 8, 8),
    (1, VF12, 0, 0, 6), (1, VF21, 0, 0, 6), (1, MA12, 4, 3, 7), (1, MA21, 4, 3, 7),
    (1, VA, 0, 0, 8),
    (2, VF12, 1, 1, 7), (2, VF21, 1, 1, 7), (2, MA12, 0, 0, 7), (2, MA21, 0, 0, 7),
    (2, VA, 0, 0, 8),
    (3, VF12, 1, 0, 7), (3, VF21, 1, 0, 7), (3, MA12, 0, 0, 7), (3, MA21, 0, 0, 7),
    (3, VA, 5, 5, 13),
    (4, VF12, 0, 0, 7), (4, VF21, 0, 0, 7), (4, MA12, 0, 0, 7), (4, MA21, 0, 0, 7),
    (4, VA, 0, 0, 13),
]
# per iteration, per op (add_ra order): (raw output before dedup, output newt after)
ANALYSIS_OPS = [
    [(8, 6), (8, 6), (10, 8), (10, 8), (2, 6), (0, 0), (0, 0), (6, 6), (6, 12), (20, 8),
     (6, 6), (10, 8), (6, 6), (0, 0)],
    [(0, 0), (0, 0), (0, 0), (0, 0), (0, 0), (4, 4), (4, 4), (0, 0), (0, 0), (0, 0),
     (0, 0), (0, 0), (0, 0), (4, 4)],
    [(0, 0), (0, 0), (0, 0), (0, 0), (1, 1), (0, 0), (0, 0), (0, 0), (1, 1), (0, 0),
     (0, 0), (0, 0), (1, 1), (0, 0)],
    [(0, 0), (1, 1), (3, 3), (3, 5), (0, 1), (0, 0), (0, 0), (1, 1), (0, 1), (3, 5),
     (1, 1), (3, 5), (1, 1), (0, 0)],
    [(0, 0)] * 14,
]
VF_FINAL = {(1, 1), (1, 2), (2, 2), (3, 3), (5, 3), (5, 4), (5, 5)}
MA_FINAL = {(1, 1), (2, 2), (3, 3), (3, 4), (4, 3), (4, 4), (5, 5)}
VA_FINAL = {(1, 1), (1, 2), (2, 1), (2, 2), (3, 3), (3, 4), (3, 5), (4, 3), (4, 4), (4, 5),
            (5, 3), (5, 4), (5, 5)}


def pairs(keys):
    a, b = gdl.unpack(keys)
    return set(zip(a.tolist(), b.tolist()))


def test_oracle_cspa_trace_is_the_hand_derived_one():
    r = oracle.cspa(ASSIGN, DEREF)
    assert r["init_iterations"] == 3 and r["iterations"] == 5
    assert [tuple(map(int, t)) for t in r["trace_init"]] == INIT_TRACE
    assert [tuple(map(int, t)) for t in r["trace"]] == ANALYSIS_TRACE
    ops = [[None] * 14 for _ in range(5)]
    for it, op, ran, raw, newt in r["oplog"]:
        assert ran == 1
        ops[it][op] = (int(raw), int(newt))
    assert ops == ANALYSIS_OPS
    assert pairs(r["extra"][0]) == VF_FINAL
    assert pairs(r["extra"][1]) == MA_FINAL
    assert pairs(r["out"]) == VA_FINAL
    assert r["derived"] == 6 + 6 + 4 + 4 + (8 + 3 + 3 + 1 + 1 + 5)


def test_hand_derived_final_sets_are_the_datalog_fixpoint():
    from test_oracle import cspa_naive
    VF, MA, VA_ = cspa_naive({(1, 2), (5, 3)}, {(1, 3), (2, 4)})
    assert (VF, MA, VA_) == (VF_FINAL, MA_FINAL, VA_FINAL)


# ------------------------------------------------- quirks CSPA cannot show
#
# In CSPA every Copy/ACopy destination has no other writer, so the zeroing of
# a destination newt by an empty source is not observable there; these LIEs
# make it (and the empty-inner zeroing of an outer delta) observable.

_HOOK2 = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_uint64),
                          ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64))
_HOOK1 = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_uint64),
                          ctypes.POINTER(ctypes.c_uint64))


def _lie_lib():
    L = oracle.lib()
    vp, u64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)
    L.orc_rel_new.restype = vp
    L.orc_rel_new.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_rel_load.argtypes = [vp, u64p, ctypes.c_size_t]
    L.orc_rel_size.restype = ctypes.c_size_t
    L.orc_rel_size.argtypes = [vp, ctypes.c_int]
    L.orc_rel_free.argtypes = [vp]
    L.orc_lie_new.restype = vp
    L.orc_lie_free.argtypes = [vp]
    L.orc_lie_add_relation.argtypes = [vp, vp, ctypes.c_int]
    L.orc_lie_add_join.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.c_int, vp, _HOOK2,
                                   ctypes.c_void_p]
    L.orc_lie_add_copy.argtypes = [vp, vp, ctypes.c_int, vp, _HOOK1]
    L.orc_lie_run.restype = ctypes.c_int
    L.orc_lie_run.argtypes = [vp]
    L.orc_lie_trace.restype = ctypes.c_size_t
    L.orc_lie_trace.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(oracle.TraceRec))]
    return L


def _load(L, name, keys):
    r = L.orc_rel_new(name.encode(), 2, 1, 0)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    L.orc_rel_load(r, k.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), k.size)
    return r


@_HOOK2
def _swap_join(inner, outer, out):  # out = (outer[1], inner[1])
    out[0] = outer[1]
    out[1] = inner[1]


@_HOOK1
def _ident(src, dst):
    dst[0] = src[0]
    dst[1] = src[1]


def _trace(L, lie):
    p = ctypes.POINTER(oracle.TraceRec)()
    n = L.orc_lie_trace(lie, ctypes.byref(p))
    return [(p[i].iteration, p[i].rel, p[i].newt, p[i].delta, p[i].full) for i in range(n)]


DELTA, FULL = 0, 1  # RelationVersion (relation.cuh:211)


def _run_lie(L, statics, updates, ops):
    """Builds a LIE over freshly loaded relations, runs it, returns
    (iterations, trace).  statics/updates: [(name, keys)]; ops: callables
    (lie, rels) -> None that add the RA ops."""
    rels = {n: _load(L, n, k) for n, k in statics + updates}
    lie = L.orc_lie_new()
    try:
        for n, _ in statics:
            L.orc_lie_add_relation(lie, rels[n], 1)
        for n, _ in updates:
            L.orc_lie_add_relation(lie, rels[n], 0)
        for add in ops:
            add(lie, rels)
        it = L.orc_lie_run(lie)
        return it, _trace(L, lie)
    finally:
        L.orc_lie_free(lie)
        for r in rels.values():
            L.orc_rel_free(r)


def test_copy_from_an_empty_source_zeroes_the_destination_newt():
    # R: update, FULL {(2,1)}; E: static {(2,3)}; S: update, empty.
    # op0: join E FULL (inner, keyed by column 0) with R DELTA (outer, keyed by
    #      its column 0 = 2), output (outer[1], inner[1]) = (1,3) into R's newt.
    # op1: Copy S DELTA -> R: S's delta is empty, so R's newt count is zeroed
    #      (copy.cu:23-26) and (1,3) is never derived.
    L = _lie_lib()
    join = lambda lie, r: L.orc_lie_add_join(lie, r["E"], FULL, r["R"], DELTA, r["R"],
                                             _swap_join, None)
    copy = lambda lie, r: L.orc_lie_add_copy(lie, r["S"], DELTA, r["R"], _ident)
    statics = [("E", gdl.pack([2], [3]))]
    updates = [("R", gdl.pack([2], [1])), ("S", np.zeros(0, np.uint64))]
    # the join alone: it0 derives (1,3); it1's delta (1,3) matches nothing in E
    assert _run_lie(L, statics, updates, [join]) == (
        2, [(0, 0, 1, 1, 2), (0, 1, 0, 0, 0), (1, 0, 0, 0, 2), (1, 1, 0, 0, 0)])
    # join then the copy from the empty delta: nothing is derived
    assert _run_lie(L, statics, updates, [join, copy]) == (
        1, [(0, 0, 0, 0, 1), (0, 1, 0, 0, 0)])
    # the copy first: it zeroes an already empty newt, then the join derives (1,3)
    assert _run_lie(L, statics, updates, [copy, join])[1][0] == (0, 0, 1, 1, 2)


def test_empty_inner_zeroes_the_outer_delta():
    # R: update {(2,1)}; Z: static, empty; E: static {(2,3)}.
    # op0: join Z FULL (empty inner) with R DELTA -> R: returns and zeroes R's
    #      delta count (join.cu:42-45).
    # op1: join E FULL with R DELTA -> R: its outer is now empty, so (1,3) is
    #      not derived.  Without op0, op1 derives (1,3).
    L = _lie_lib()
    join_e = lambda lie, r: L.orc_lie_add_join(lie, r["E"], FULL, r["R"], DELTA, r["R"],
                                               _swap_join, None)
    join_z = lambda lie, r: L.orc_lie_add_join(lie, r["Z"], FULL, r["R"], DELTA, r["R"],
                                               _swap_join, None)
    statics = [("E", gdl.pack([2], [3])), ("Z", np.zeros(0, np.uint64))]
    updates = [("R", gdl.pack([2], [1]))]
    assert _run_lie(L, statics, updates, [join_e]) == (2, [(0, 0, 1, 1, 2), (1, 0, 0, 0, 2)])
    assert _run_lie(L, statics, updates, [join_z, join_e]) == (1, [(0, 0, 0, 0, 1)])
    # the other order: op0 derives (1,3) before the empty inner zeroes the delta
    assert _run_lie(L, statics, updates, [join_e, join_z])[1][0] == (0, 0, 1, 1, 2)


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["merge", "auto"])
def test_gpu_cspa_trace_is_the_hand_derived_one(ctx, mode):
    gdl.set_full_mode(gdl.FULL_MODE_MERGE if mode == "merge" else gdl.FULL_MODE_AUTO)
    try:
        got = ctx.run_cspa(ASSIGN, DEREF, out_cap=1000, extra_cap=1000)
    finally:
        gdl.set_full_mode(gdl.FULL_MODE_AUTO)
    assert got["init_iterations"] == 3 and got["iterations"] == 5
    assert [tuple(map(int, t)) for t in got["trace_init"]] == INIT_TRACE
    tr = [tuple(map(int, t)) for t in got["trace"]]
    if mode == "merge":
        assert tr == ANALYSIS_TRACE
    else:
        # ValueAlias is an append-mode relation in auto mode (its FULL is no
        # join's inner): its newt is the raw join output, deduplicated only
        # against FULL's set, so only its delta and full counts are the
        # reference's; the other four relations are unchanged
        assert [t for t in tr if t[1] != VA] == [t for t in ANALYSIS_TRACE if t[1] != VA]
        assert [(t[0], t[3], t[4]) for t in tr if t[1] == VA] == \
            [(t[0], t[3], t[4]) for t in ANALYSIS_TRACE if t[1] == VA]
        # raw VA newt per iteration: the joins' outputs (ANALYSIS_OPS raw
        # sizes of ops 2, 3, 9, 11)
        assert [t[2] for t in tr if t[1] == VA] == \
            [sum(ANALYSIS_OPS[i][o][0] for o in (2, 3, 9, 11)) for i in range(5)]
    assert pairs(got["out"]) == VA_FINAL
