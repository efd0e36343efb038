"""The Datalog front end (include/gdl_datalog.h): programs in the Souffle subset
lowered to LIEs of RA operators and run on the engine.

Parity: the reference ships its programs as Souffle sources
(test/souffle/tc.dl, sg.dl, cspa.dl) next to the hand-written RA encodings
(test/tc.cu, test/sg.cu, test/cspa.cu).  The same rules, run through the
front end, must give exactly the oracle's TC, SG and CSPA relations (the
restated hand encodings), bit for bit.  Programs with no hand encoding
(several strata, reversed static copies, copy rules inside a recursive
stratum, an input relation that rules extend, comparisons in 3-atom rules)
are checked against tests/datalog_naive.py, itself pinned here against the
oracle.  Inputs are duplicate-free: with duplicate rows the hand encodings
seed their first delta from raw rows (the oracle restates that), while a
Datalog program's result depends only on the set of input tuples.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import gdl, read_edges
from datalog_naive import evaluate

TC_LEFT = """
.decl edge(from: number, to: number)
.input edge(IO=file, deliminator="\\t")
.decl path(from: number, to: number)
.output path
path(x, y) :- edge(x, y).
path(x, y) :- path(x, z), edge(z, y).
"""
TC_RIGHT = TC_LEFT.replace("path(x, z), edge(z, y)", "edge(x, z), path(z, y)")
# test/souffle/tc.dl's form: both body atoms recursive
TC_SQUARE = TC_LEFT.replace("path(x, z), edge(z, y)", "path(x, z), path(z, y)")

# test/souffle/sg.dl
SG = """
.decl edge(x: number, y: number)
.input edge
.decl sg(x: number, y: number)
.printsize sg
sg(x, y) :- edge(p, x), edge(p, y), x != y.
sg(x, y) :- edge(a, x), sg(a, b), edge(b, y).
"""

# test/souffle/cspa.dl:25-39
CSPA = """
.decl assign(src: number, dest: number)
.input assign
.decl dereference(src: number, dest: number)
.input dereference
.decl ValueFlow(src: number, dest: number)
.output ValueFlow
.decl ValueAlias(src: number, dest: number)
.output ValueAlias
.decl MemoryAlias(src: number, dest: number)
.printsize MemoryAlias
ValueFlow(x, y) :- ValueFlow(x, z), ValueFlow(z, y).
ValueAlias(x, y) :- ValueFlow(z, x), ValueFlow(z, y).
ValueFlow(x, y) :- assign(x, z), MemoryAlias(z, y).
MemoryAlias(x, w) :- dereference(y, x), ValueAlias(y, z), dereference(z, w).
ValueAlias(x, y) :- ValueFlow(z, x), MemoryAlias(z, w), ValueFlow(w, y).
ValueFlow(y, x) :- assign(y, x).
ValueFlow(x, x) :- assign(x, y).
ValueFlow(x, x) :- assign(y, x).
MemoryAlias(x, x) :- assign(y, x).
MemoryAlias(x, x) :- assign(x, y).
"""

# several strata; reach is stored reversed (its recursive join reads it by
# column 1), so `pairs` reads it through a reversed static copy; sym is an
# input that a copy rule extends inside its own recursive stratum; odd/even
# are mutually recursive with an always-empty inner joined before a
# non-empty one on the same delta; `wide` is a 3-atom join with a comparison
# in an init LIE; `back` is a copy from an earlier stratum.
MIXED = """
/* a block comment */
.decl e(a: number, b: number)
.decl seed(a: number, b: number)
.decl sym(a: number, b: number)
.decl nothing(a: number, b: number)
.input e, seed, sym
.input nothing
.decl reach(a: number, b: number)
.decl back(a: number, b: number)
.decl pairs(a: number, b: number)
.decl odd(a: number, b: number)
.decl even(a: number, b: number)
.decl wide(a: number, b: number)
.output reach, back, pairs, sym
.output odd, even, wide
reach(x, y) :- seed(x, y).
reach(x, y) :- reach(x, z), e(z, y).
back(y, x) :- reach(x, y).
pairs(x, y) :- reach(z, x), reach(z, y), x != y.
sym(x, y) :- e(x, y).
sym(x, y) :- sym(y, x).
odd(x, y) :- e(x, y).
even(x, y) :- odd(x, z), nothing(z, y).
even(x, y) :- odd(x, z), e(z, y).
odd(x, y) :- even(x, z), e(z, y).
wide(x, y) :- e(x, z), sym(z, w), e(w, y), x < y.
"""


# single-source reachability: unary relations, facts, a unary delta joined
# with a binary inner, a reversed static read by a unary atom, a 3-atom join
# starting at a unary atom, and copies between unary and binary relations
REACH = """
.decl edge(a: number, b: number)
.input edge
.decl source(a: number)
source(0).
source(7).
.decl reach(a: number)
.decl hop2(a: number, b: number)
.decl entered(a: number)
.decl loop(a: number, b: number)
.output reach, hop2, entered, loop
reach(x) :- source(x).
reach(y) :- reach(x), edge(x, y).
hop2(s, y) :- source(s), edge(s, z), edge(z, y), s != y.
entered(x) :- reach(x), edge(_, x).
loop(x, x) :- reach(x).
"""


# constants in rule bodies: each atom with a constant becomes a selection
# relation (#s) joined from a one-fact relation (#c); `via` selects inside
# its own recursive stratum
CONSTS = """
.decl edge(a: number, b: number)
.input edge
.decl reach(a: number)
.decl hub(a: number, b: number)
.decl via(a: number, b: number)
.decl hop4(a: number, b: number)
.output reach, hub, via, hop4
reach(y) :- edge(0, y).
reach(y) :- reach(x), edge(x, y).
hub(x, y) :- edge(x, 3), edge(x, y).
via(x, y) :- edge(x, y).
via(x, y) :- via(x, z), via(z, y), via(z, 1).
hop4(x, y) :- edge(x, a), edge(a, b), edge(b, c), edge(c, y).
"""


# comparisons with ids and in one-atom rules (a copy's filter); `far`
# runs in the linear loop with its comparison against an id
COMPARE = """
.decl edge(a: number, b: number)
.input edge
.decl low(a: number, b: number)
.decl loop(a: number, b: number)
.decl far(a: number, b: number)
.decl mid(a: number)
.output low, loop, far, mid
low(x, y) :- edge(x, y), y < 30.
loop(x, y) :- edge(x, y), x = y.
far(x, y) :- edge(x, y).
far(x, y) :- far(x, z), edge(z, y), 20 <= y.
mid(y) :- edge(x, y), far(y, z), 10 > y.
"""


def uniq(keys):
    return np.unique(np.asarray(keys, dtype=np.uint64))


def mixed_inputs(seed):
    rng = np.random.default_rng(seed)
    e = uniq(gdl.gen_er(seed, 60, 110))
    s = rng.integers(0, 60, size=(8, 2))
    sy = rng.integers(0, 60, size=(5, 2))
    return dict(e=e, seed=uniq(gdl.pack(s[:, 0], s[:, 1])),
                sym=uniq(gdl.pack(sy[:, 0], sy[:, 1])), nothing=np.zeros(0, dtype=np.uint64))


# ---- CPU: the checker, plans, rejections ---------------------------------------

@pytest.mark.parametrize("src", [TC_LEFT, TC_RIGHT, TC_SQUARE])
def test_naive_evaluator_matches_the_oracle_tc(src):
    e = uniq(gdl.gen_er(7, 80, 160))
    assert np.array_equal(evaluate(src, {"edge": e})["path"], oracle.tc(e)["out"])


def test_naive_evaluator_matches_the_oracle_sg():
    e = uniq(gdl.gen_rmat(3, 6, 120))
    assert np.array_equal(evaluate(SG, {"edge": e})["sg"], oracle.sg(e)["out"])


def test_naive_evaluator_matches_the_oracle_cspa():
    a, d = gdl.gen_cspa(9, 40, 1.1, 50, 25)
    a, d = uniq(a), uniq(d)
    got = evaluate(CSPA, {"assign": a, "dereference": d})
    want = oracle.cspa(a, d)
    assert np.array_equal(got["ValueAlias"], want["out"])
    assert np.array_equal(got["ValueFlow"], want["extra"][0])
    assert np.array_equal(got["MemoryAlias"], want["extra"][1])


def test_naive_evaluator_matches_bfs_on_unary_relations():
    e = uniq(gdl.gen_er(4, 300, 420))
    src, dst = (e >> np.uint64(32)).astype(np.int64), (e & np.uint64(0xffffffff)).astype(np.int64)
    seen, frontier = {0, 7}, [0, 7]
    while frontier:
        frontier = [int(d) for s_ in frontier for d in dst[src == s_] if int(d) not in seen]
        frontier = sorted(set(frontier))
        seen.update(frontier)
    got = evaluate(REACH, {"edge": e})
    assert np.array_equal(got["reach"], np.array(sorted(x << 32 for x in seen), dtype=np.uint64))
    assert np.array_equal(got["loop"], np.array(sorted((x << 32) | x for x in seen), dtype=np.uint64))


def test_unary_plan():
    p = gdl.Datalog(REACH)
    assert p.arity("reach") == 1 and p.arity("edge") == 2 and p.arity("nope") == 0
    plan = p.plan
    assert "join edge_2__1_2 FULL x reach_1__1 DELTA -> reach_1__1 (i1)" in plan
    assert "copy source_1__1 FULL -> reach_1__1 (i0)" in plan
    assert "copy reach_1__1 FULL -> loop_2__1_2 (i0, i0)" in plan
    assert "join edge_2__2_1 FULL x reach_1__1 FULL -> entered_1__1 (o0)" in plan
    assert p.kept() == ["reach", "hop2", "entered", "loop"]


def test_body_constants_become_selections():
    p = gdl.Datalog(CONSTS)
    assert p.kept() == ["reach", "hub", "via", "hop4"]
    plan = p.plan
    assert "join edge_2__1_2 FULL x #c0_1__1 FULL -> #s_edge_0_0_1__1 (i1)" in plan
    assert "copy #s_edge_0_0_1__1 FULL -> reach_1__1 (i0)" in plan
    # via(z, 1) reads via: the selection joins via's recursive stratum
    assert any(l.startswith("lie via, #s_via_1_1") or l.startswith("lie #s_via_1_1, via")
               for l in plan.splitlines())
    e = uniq(gdl.gen_er(5, 25, 45))
    got = evaluate(CONSTS, {"edge": e})
    k = e[(e >> np.uint64(32)) == 0] & np.uint64(0xffffffff)
    assert set(k.tolist()) <= set((got["reach"] >> np.uint64(32)).tolist())


def lie_blocks(plan):
    """[(header line, [operator lines])] of a plan."""
    out = []
    for line in plan.splitlines():
        if line.startswith("lie "):
            out.append((line, []))
        elif line.startswith("  "):
            out[-1][1].append(line.strip())
    return out


def test_linear_tc_lowers_to_one_join_in_the_stored_order():
    left = lie_blocks(gdl.Datalog(TC_LEFT).plan)
    assert [len(ops) for _, ops in left] == [1, 1]
    assert "update: path_2__2_1" in left[1][0]  # path(x, z) is read by z: stored reversed
    assert left[1][1] == ["join edge_2__1_2 FULL x path_2__2_1 DELTA -> path_2__2_1 (i1, o1)"]
    right = gdl.Datalog(TC_RIGHT).plan
    assert "build edge_2__2_1 from edge_2__1_2" in right
    assert "join edge_2__2_1 FULL x path_2__1_2 DELTA -> path_2__1_2 (i1, o1)" in right


def test_square_tc_keeps_a_mirror():
    blocks = lie_blocks(gdl.Datalog(TC_SQUARE).plan)
    rec = blocks[1][1]
    assert sum(op.startswith("join") for op in rec) == 2
    assert rec[-1] == "acopy path_2__1_2 NEWT -> path_2__2_1 (i1, i0)"


def test_sg_lowers_to_the_hand_encoding():
    blocks = lie_blocks(gdl.Datalog(SG).plan)
    assert blocks[0][1] == ["join edge_2__1_2 FULL x edge_2__1_2 FULL -> sg_2__1_2 (o1, i1) "
                            "if c0 != c1"]
    tmp = blocks[1][1][0].split("-> ")[1].split()[0]
    assert blocks[1][1] == [f"join edge_2__1_2 FULL x sg_2__1_2 DELTA -> {tmp} (o1, i1)",
                            f"join edge_2__1_2 FULL x {tmp} NEWT -> sg_2__1_2 (o1, i1)"]


def test_cspa_plan_shape():
    p = gdl.Datalog(CSPA)
    blocks = lie_blocks(p.plan)
    assert len(blocks) == 2
    init, rec = blocks[0][1], blocks[1][1]
    assert len(init) == 5 and all(op.startswith("copy assign_2__1_2 FULL") for op in init)
    assert sum(op.startswith("join") for op in rec) == 13
    assert [op for op in rec if op.startswith("acopy")] == [
        "acopy ValueFlow_2__1_2 NEWT -> ValueFlow_2__2_1 (i1, i0)",
        "acopy MemoryAlias_2__1_2 NEWT -> MemoryAlias_2__2_1 (i1, i0)"]
    assert p.kept() == ["ValueFlow", "ValueAlias", "MemoryAlias"]


def test_strata_run_in_dependency_order():
    plan = gdl.Datalog(MIXED).plan
    order = [line.split(":")[0][4:] for line in plan.splitlines() if line.startswith("lie ")]
    assert order.index("reach") < order.index("back") < order.index("pairs")
    assert order.index("sym") < order.index("wide")
    assert "odd, even" in order or "even, odd" in order
    assert "build reach_2__1_2 from reach_2__2_1" in plan  # pairs reads reach by column 0


def test_kept_defaults_to_every_derived_relation():
    src = TC_LEFT.replace(".output path\n", "") + ".decl hop(a: number, b: number)\n" \
        "hop(x, y) :- path(x, z), edge(z, y).\n"
    assert gdl.Datalog(src).kept() == ["path", "hop"]


BAD = [
    ("path(x, y) :- edge(x, y).", gdl.GDL_ERR_INVALID_ARG, "relation path is not declared"),
    (".decl r(a: number, b: number, c: number)", 4, "arity 3"),
    (".decl r(a: symbol, b: number)", 4, "attribute type symbol"),
    (".decl r(a: number, b: number)\n.decl r(a: number, b: number)", gdl.GDL_ERR_INVALID_ARG,
     "declared twice"),
    (".decl r(a: number, b: number)\nr(x, 1) :- r(x, y).", 4, "constants in rule heads"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(1, 2), r(x, y).", 4, "a ground atom in a rule body"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, y), r(1, _).", 4,
     "holding a constant and no variable"),
    (".decl r(a: number)\nr(x) :- r(x), r(1).", 4, "a ground atom in a rule body"),
    (".decl r(a: number, b: number)\nr(1, x).", 4, "constants in rule heads"),
    (".decl r(a: number, b: number)\nr(x, y).", gdl.GDL_ERR_INVALID_ARG, "a fact with variables"),
    (".decl r(a: number, b: number)\nr(4294967296, 1).", gdl.GDL_ERR_INVALID_ARG, "bad id"),
    (".decl r(a: number)\nr(1, 2).", gdl.GDL_ERR_INVALID_ARG, "r takes 1 argument, not 2"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, y), !r(y, x).", 4, "negation"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, z), r(w, y).", 4, "cross product"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, y), r(y, x).", 4, "two variables"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, x), r(x, y).", 4, "twice in one atom"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, z)", gdl.GDL_ERR_INVALID_ARG, "line 2"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, z).", gdl.GDL_ERR_INVALID_ARG,
     "head variable y is not bound"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, y), 2 != 3.", 4, "a comparison of two ids"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, z), r(z, y), z != y.", 4,
     "not a head variable"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, a), r(a, b), r(b, c), r(c, d), r(d, y).",
     4, "more than four"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, z), r(z, y); r(y, x).", 4, "disjunction"),
    (".decl r(a: number, b: number)\n.type T = number", 4, "directive .type"),
    (".decl r(a: number, b: number)\nr(x, y) :- r(x, z), r(z, w), r(w, y), r(z, y).", 4,
     "three variables live"),
    ("/* open", gdl.GDL_ERR_INVALID_ARG, "unterminated comment"),
]


@pytest.mark.parametrize("src,status,fragment", BAD)
def test_unsupported_and_invalid_programs_are_rejected(src, status, fragment):
    with pytest.raises(gdl.GdlError) as e:
        gdl.Datalog(src)
    assert e.value.status == status
    assert fragment in str(e.value)


# ---- GPU: parity --------------------------------------------------------------

TC_INPUTS = {
    "er_500": lambda: gdl.gen_er(3, 500, 1500),
    "tree_5k": lambda: gdl.gen_tree(11, 5000, 100),
    "data_7035": lambda: read_edges("data_7035.txt"),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(TC_INPUTS))
@pytest.mark.parametrize("form", ["left", "right", "square"])
def test_tc_programs_match_the_oracle(ctx, form, name):
    e = uniq(TC_INPUTS[name]())
    src = {"left": TC_LEFT, "right": TC_RIGHT, "square": TC_SQUARE}[form]
    prog = gdl.Datalog(src)
    got = prog.run(ctx, edge=e)
    want = oracle.tc(e)
    assert np.array_equal(got["path"], want["out"])
    derived, its = prog.stats()
    assert derived == want["out"].size and len(its) == 2


@pytest.mark.gpu
@pytest.mark.parametrize("name,make", [("rmat_10", lambda: gdl.gen_rmat(6, 10, 3000)),
                                       ("tree_2k", lambda: gdl.gen_tree(2, 2000, 30))])
def test_sg_program_matches_the_oracle(ctx, name, make):
    e = uniq(make())
    got = gdl.Datalog(SG).run(ctx, edge=e)
    assert np.array_equal(got["sg"], oracle.sg(e)["out"])


@pytest.mark.gpu
@pytest.mark.parametrize("seed,n_ids,na,nd", [(5, 150, 200, 100), (8, 400, 600, 300)])
def test_cspa_program_matches_the_oracle(ctx, seed, n_ids, na, nd):
    a, d = gdl.gen_cspa(seed, n_ids, 1.1, na, nd)
    a, d = uniq(a), uniq(d)
    got = gdl.Datalog(CSPA).run(ctx, assign=a, dereference=d)
    want = oracle.cspa(a, d)
    assert np.array_equal(got["ValueAlias"], want["out"])
    assert np.array_equal(got["ValueFlow"], want["extra"][0])
    assert np.array_equal(got["MemoryAlias"], want["extra"][1])


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mixed_program_matches_the_naive_evaluator(ctx, seed):
    inputs = mixed_inputs(seed)
    want = evaluate(MIXED, inputs)
    got = gdl.Datalog(MIXED).run(ctx, inputs)
    assert sorted(got) == sorted(want)
    for rel in want:
        assert np.array_equal(got[rel], want[rel]), rel
    assert got["even"].size > 0  # the empty inner did not starve the other join


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [4, 5])
def test_reach_program_matches_the_naive_evaluator(ctx, seed):
    e = uniq(gdl.gen_er(seed, 300, 420))
    want = evaluate(REACH, {"edge": e})
    prog = gdl.Datalog(REACH)
    got = prog.run(ctx, edge=e)
    for rel in want:
        assert np.array_equal(got[rel], want[rel]), rel
    assert got["reach"].size > 2


def test_comparison_plans():
    plan = gdl.Datalog(COMPARE).plan
    assert "copy edge_2__1_2 FULL -> low_2__1_2 (i0, i1) if c1 < 30" in plan
    assert "copy edge_2__1_2 FULL -> loop_2__1_2 (i0, i1) if c0 = c1" in plan
    assert "-> far_2__2_1 (i1, o1) if c0 >= 20" in plan  # y <= ... flipped, far stored reversed


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [3, 4])
def test_comparison_program_matches_the_naive_evaluator(ctx, seed):
    e = uniq(np.concatenate([gdl.gen_er(seed, 50, 140), gdl.pack(np.arange(5), np.arange(5))]))
    want = evaluate(COMPARE, {"edge": e})
    got = gdl.Datalog(COMPARE).run(ctx, edge=e)
    for rel in want:
        assert np.array_equal(got[rel], want[rel]), rel
    assert all(got[r].size > 0 for r in want)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [6, 9])
def test_constants_program_matches_the_naive_evaluator(ctx, seed):
    e = uniq(gdl.gen_er(seed, 25, 45))
    want = evaluate(CONSTS, {"edge": e})
    got = gdl.Datalog(CONSTS).run(ctx, edge=e)
    for rel in want:
        assert np.array_equal(got[rel], want[rel]), rel
    assert got["reach"].size > 0 and got["via"].size > 0


@pytest.mark.gpu
def test_unary_inputs_and_facts_merge(ctx):
    src = (".decl s(a: number)\n.input s\n.decl e(a: number, b: number)\n.input e\n"
           ".decl r(a: number)\n.output r\ns(5).\nr(x) :- s(x).\nr(y) :- r(x), e(x, y).\n")
    prog = gdl.Datalog(src)
    e = gdl.pack(np.array([1, 2, 5, 9]), np.array([2, 3, 9, 9]))
    got = prog.run(ctx, s=np.array([1 << 32], dtype=np.uint64), e=e)["r"]
    assert (got >> np.uint64(32)).tolist() == [1, 2, 3, 5, 9]
    with pytest.raises(gdl.GdlError) as err:
        prog.run(ctx, s=np.array([(1 << 32) | 1], dtype=np.uint64), e=e)
    assert "nonzero column 1" in str(err.value)


@pytest.mark.gpu
def test_runs_replace_results_and_reject_bad_inputs(ctx):
    prog = gdl.Datalog(TC_LEFT)
    e1 = uniq(gdl.gen_er(1, 200, 400))
    e2 = uniq(gdl.gen_er(2, 100, 150))
    assert np.array_equal(prog.run(ctx, edge=e1)["path"], oracle.tc(e1)["out"])
    assert np.array_equal(prog.run(ctx, edge=e2)["path"], oracle.tc(e2)["out"])
    # duplicates and order of the input rows do not matter
    shuffled = np.concatenate([e2[::-1], e2[:10]])
    assert np.array_equal(prog.run(ctx, edge=shuffled)["path"], oracle.tc(e2)["out"])
    assert prog.run(ctx, edge=np.zeros(0, dtype=np.uint64))["path"].size == 0
    with pytest.raises(gdl.GdlError) as e:
        prog.run(ctx)
    assert e.value.status == gdl.GDL_ERR_INVALID_ARG and "no tuples given" in str(e.value)
    with pytest.raises(gdl.GdlError) as e:
        prog.run(ctx, edge=e1, path=e1)
    assert "not an .input relation" in str(e.value)
    with pytest.raises(gdl.GdlError):
        prog.result("edge")


# ---- Slog RA plans ---------------------------------------------------------------

REF_CSPA_SLOGC = "/root/reference/test/souffle/cspa.slogc"

# TC as a Slog compiler's plan, written for these tests in that format:
# path(x, y) :- edge(x, y).   path(x, y) :- path(x, z), edge(z, y).
# A version stores (index columns, id, other columns); the join projects
# [z, path.id, x, edge.id, y] onto (2 4).
TC_SLOGC = """
(slog-prog
 ((relation-decl rel__edge__2__1__2 edge 2 #t 2 (1 2))
  (relation-decl rel__edge__2__1 edge 1 #f 2 (1))
  (relation-decl rel__path__2__1__2 path 2 #t 2 (1 2))
  (relation-decl rel__path__2__2 path 1 #f 2 (2)))
 ((scc-decl scc0 0 #f
   ((scc-rel rel__edge__2__1__2 #f #f) (scc-rel rel__path__2__1__2 #t #f))
   ((copy rel__path__2__1__2 rel__edge__2__1__2 FULL (0 1))))
  (scc-decl scc1 1 #f
   ((scc-rel rel__edge__2__1 #t #f) (scc-rel rel__edge__2__1__2 #t #f))
   ((acopy rel__edge__2__1 rel__edge__2__1__2 DELTA (0 2 1))))
  (scc-decl scc2 2 #t
   ((scc-rel rel__path__2__1__2 #t #f) (scc-rel rel__path__2__2 #t #f)
    (scc-rel rel__edge__2__1 #f #f))
   ((join rel__path__2__1__2 rel__path__2__2 DELTA rel__edge__2__1 FULL (2 4))
    (acopy rel__path__2__2 rel__path__2__1__2 DELTA (1 2 0)))))
 ((0 2) (1 2)))
"""


def test_slog_tc_plan_reads_back_to_tc():
    p = gdl.Datalog(TC_SLOGC, slog=True)
    assert "path(v1, v2) :- path(v1, v0), edge(v0, v2)." in p.source
    assert ".input edge" in p.source and ".output path" in p.source
    assert lie_blocks(p.plan)[1][1] == [
        "join edge_2__1_2 FULL x path_2__2_1 DELTA -> path_2__2_1 (i1, o1)"]
    e = uniq(gdl.gen_er(7, 80, 160))
    assert np.array_equal(evaluate(p.source, {"edge": e})["path"], oracle.tc(e)["out"])


@pytest.mark.skipif(not os.path.exists(REF_CSPA_SLOGC),
                    reason="the reference's Slog plan is only present where it is checked out")
def test_reference_cspa_slog_plan_reads_back_to_cspa():
    with open(REF_CSPA_SLOGC) as f:
        p = gdl.Datalog(f.read(), slog=True)
    assert p.kept() == ["value_alias", "value_flow", "memory_alias"]
    a, d = gdl.gen_cspa(9, 40, 1.1, 50, 25)
    a, d = uniq(a), uniq(d)
    got = evaluate(p.source, {"assign": a, "dereference": d})
    want = oracle.cspa(a, d)
    assert np.array_equal(got["value_alias"], want["out"])
    assert np.array_equal(got["value_flow"], want["extra"][0])
    assert np.array_equal(got["memory_alias"], want["extra"][1])


SLOG_BAD = [
    ("(not-a-plan)", gdl.GDL_ERR_INVALID_ARG, "not a slog-prog plan"),
    ("(slog-prog (", gdl.GDL_ERR_INVALID_ARG, "unbalanced"),
    (TC_SLOGC.replace("(2 4)", "(1 4)"), 4, "provenance id in a rule head"),
    (TC_SLOGC.replace("(2 4)", "(2 9)"), gdl.GDL_ERR_INVALID_ARG, "out of range"),
    (TC_SLOGC.replace("rel__edge__2__1 FULL", "rel__edge__2__1__2 FULL"), 4,
     "different join columns"),
    (TC_SLOGC.replace("(copy rel__path__2__1__2", "(copy rel__path__2__2"), 4,
     "into an index version"),
    (TC_SLOGC.replace("edge 2 #t 2 (1 2)", "edge 3 #t 3 (1 2 3)"), 4, "arity 3"),
    (TC_SLOGC.replace("(acopy rel__edge__2__1 rel__edge__2__1__2",
                      "(acopy rel__edge__2__1 rel__path__2__1__2"), 4, "acopy between two"),
    (TC_SLOGC.replace("(2 4)", "(2 99999999999999999999)"), gdl.GDL_ERR_INVALID_ARG,
     "expected a small number"),
    ("(" * 100000 + ")" * 100000, gdl.GDL_ERR_INVALID_ARG, "nests too deeply"),
]


@pytest.mark.parametrize("plan,status,fragment", SLOG_BAD)
def test_bad_slog_plans_are_rejected(plan, status, fragment):
    with pytest.raises(gdl.GdlError) as e:
        gdl.Datalog(plan, slog=True)
    assert e.value.status == status
    assert fragment in str(e.value)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["er_500", "data_7035"])
def test_slog_tc_plan_matches_the_oracle(ctx, name):
    e = uniq(TC_INPUTS[name]())
    got = gdl.Datalog(TC_SLOGC, slog=True).run(ctx, edge=e)
    assert np.array_equal(got["path"], oracle.tc(e)["out"])


REF_SOUFFLE = "/root/reference/test/souffle"


@pytest.mark.skipif(not os.path.isdir(REF_SOUFFLE),
                    reason="the reference's Souffle programs are only present where it is checked out")
def test_the_reference_souffle_programs_compile_as_written():
    def prog(name):
        with open(os.path.join(REF_SOUFFLE, name)) as f:
            return f.read()
    e = uniq(gdl.gen_er(7, 80, 160))
    tc = gdl.Datalog(prog("tc.dl"))
    assert tc.kept() == ["path"]
    assert np.array_equal(evaluate(tc.source, {"edge": e})["path"], oracle.tc(e)["out"])
    sg = gdl.Datalog(prog("sg.dl"))
    g = uniq(gdl.gen_rmat(3, 6, 120))
    assert np.array_equal(evaluate(sg.source, {"edge": g})["sg"], oracle.sg(g)["out"])
    cspa = gdl.Datalog(prog("cspa.dl"))
    assert cspa.kept() == ["ValueFlow", "ValueAlias", "MemoryAlias", "tmp"]
    a, d = gdl.gen_cspa(9, 40, 1.1, 50, 25)
    a, d = uniq(a), uniq(d)
    got = evaluate(cspa.source, {"assign": a, "dereference": d})
    want = oracle.cspa(a, d)
    assert np.array_equal(got["ValueAlias"], want["out"])
    assert np.array_equal(got["ValueFlow"], want["extra"][0])
    # the programs outside the subset (symbols, choice domains, arity 3) are refused
    for name in ["bip.dl", "spanning.dl", "choice_total.dl", "path_length.dl"]:
        with pytest.raises(gdl.GdlError) as err:
            gdl.Datalog(prog(name))
        assert err.value.status == 4, name


# ---- the reference's data files through the front end ------------------------------

def read_dl(name):
    path = os.path.join(os.path.dirname(gdl.__file__), "programs", "dl", name)
    with open(path) as f:
        return f.read()


@pytest.mark.gpu
@pytest.mark.parametrize("fname,size", [("data_48232.txt", 14_503_742),
                                        ("data_26013.txt", 62_056_583)])
def test_sg_dl_sizes_match_cudf(ctx, fname, size):
    # the sizes cuDF published for these graphs (test/cuDF/sg.json: cti, p2p-Gnutella09),
    # an implementation independent of the reference's engine
    got = gdl.Datalog(read_dl("sg.dl")).run(ctx, edge=read_edges(fname))
    assert got["sg"].size == size
    assert np.all(got["sg"][1:] > got["sg"][:-1])


@pytest.mark.gpu
@pytest.mark.parametrize("fname", ["data_5.txt", "hpc_talk.txt", "data_7035.txt",
                                   "data_21693.txt", "data_23874.txt"])
def test_tc_dl_on_the_data_files_matches_the_oracle(ctx, fname):
    e = uniq(read_edges(fname))
    got = gdl.Datalog(read_dl("tc.dl")).run(ctx, edge=e)
    assert np.array_equal(got["path"], oracle.tc(e)["out"])


@pytest.mark.gpu
def test_raw_inputs_with_duplicate_rows_where_the_inputs_are_static(ctx):
    # SG's edge and CSPA's assign / dereference are static relations, so the
    # hand encodings' raw-row seeding of the first delta does not touch them:
    # the front end gives the oracle's relations on the raw rows as well
    e = read_edges("data_7035.txt")  # 6 duplicate rows
    assert e.size > np.unique(e).size
    assert np.array_equal(gdl.Datalog(SG).run(ctx, edge=e)["sg"], oracle.sg(e)["out"])
    a, d = gdl.gen_cspa(5, 150, 1.1, 200, 100)
    a, d = np.concatenate([a, a[:20]]), np.concatenate([d, d[:10]])  # repeated rows
    got = gdl.Datalog(CSPA).run(ctx, assign=a, dereference=d)
    want = oracle.cspa(a, d)
    assert np.array_equal(got["ValueAlias"], want["out"])
    assert np.array_equal(got["ValueFlow"], want["extra"][0])
    assert np.array_equal(got["MemoryAlias"], want["extra"][1])


@pytest.mark.gpu
@pytest.mark.parametrize("stop", [1, 4])
def test_a_stopped_datalog_run_resumes_to_the_same_relations(ctx, tmp_path, stop):
    # the recursive CSPA LIE (mirrors, tmps) stops right after its snapshot at
    # iteration `stop`; the next run resumes every LIE from its snapshot
    a, d = gdl.gen_cspa(8, 400, 1.1, 600, 300)
    a, d = uniq(a), uniq(d)
    prog = gdl.Datalog(CSPA)
    want = prog.run(ctx, assign=a, dereference=d)
    try:
        gdl.set_checkpoint(str(tmp_path), every=0, resume=False, stop_after=stop)
        with pytest.raises(gdl.GdlError) as err:
            prog.run(ctx, assign=a, dereference=d)
        assert err.value.status == gdl.GDL_ERR_BUSY
        gdl.set_checkpoint(str(tmp_path), every=0, resume=True, stop_after=-1)
        got = prog.run(ctx, assign=a, dereference=d)
    finally:
        gdl.set_checkpoint(None)
    for rel in want:
        assert np.array_equal(got[rel], want[rel]), rel
