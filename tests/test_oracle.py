"""Pins the CPU oracle (oracle/gdl_oracle.c) before it is trusted as the checker.

Each check comes from outside the oracle's own code: Python's sorted()/set
algebra for the primitives, the reference's worked TC examples
(data/tc.png, data/JOIN.pdf), shortest-distance closed forms (scipy BFS),
cuDF's published SG sizes (test/cuDF/sg.json), a scipy boolean-matrix SG
fixpoint, and a naive Datalog evaluation of test/souffle/cspa.dl.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

import oracle
from conftest import DATA, gdl, read_edges

RNG = np.random.default_rng(12345)


def keys_of(pairs):
    return sorted((int(a) << 32) | int(b) for a, b in pairs)


def pairs_of(keys):
    return {(int(k) >> 32, int(k) & 0xffffffff) for k in keys}


# ------------------------------------------------------------------ primitives

@pytest.mark.parametrize("case", ["random", "dups", "equal", "sorted", "reversed", "maxid",
                                  "empty", "one"])
def test_sort_unique_matches_python(case):
    n = 5000
    if case == "random":
        a = RNG.integers(0, 1 << 62, n, dtype=np.uint64)
    elif case == "dups":
        a = RNG.integers(0, 50, n, dtype=np.uint64)
    elif case == "equal":
        a = np.full(n, 77, dtype=np.uint64)
    elif case == "sorted":
        a = np.arange(n, dtype=np.uint64) * np.uint64(3)
    elif case == "reversed":
        a = (np.arange(n, dtype=np.uint64) * np.uint64(3))[::-1].copy()
    elif case == "maxid":
        a = gdl.pack(RNG.integers(0xfffffff0, 0xffffffff, n, endpoint=True),
                     RNG.integers(0xfffffff0, 0xffffffff, n, endpoint=True))
    elif case == "empty":
        a = np.zeros(0, dtype=np.uint64)
    else:
        a = np.array([5], dtype=np.uint64)
    got = oracle.sort_unique(a)
    assert got.tolist() == sorted(set(int(x) for x in a))


def test_merge_union_and_set_difference_match_python():
    for trial in range(20):
        a = np.unique(RNG.integers(0, 3000, RNG.integers(0, 2000), dtype=np.uint64))
        b = np.unique(RNG.integers(0, 3000, RNG.integers(0, 2000), dtype=np.uint64))
        sa, sb = set(a.tolist()), set(b.tolist())
        assert oracle.merge_union(a, b).tolist() == sorted(sa | sb)
        assert oracle.set_difference(b, a).tolist() == sorted(sb - sa)


# ------------------------------------------------------------------ TC

def test_tc_kat_hpc_talk():
    # data/tc.png: 9 tuples, per-iteration new tuples [3, 1, 0], 3 iterations
    r = oracle.tc(read_edges("hpc_talk.txt"))
    assert r["n_out"] == 9
    assert [t[3] for t in r["trace"]] == [3, 1, 0]
    assert r["iterations"] == 3
    assert pairs_of(r["out"]) == {(1, 2), (1, 3), (2, 4), (3, 4), (4, 5), (1, 4), (2, 5),
                                  (3, 5), (1, 5)}


def test_tc_kat_join_pdf():
    # data/JOIN.pdf on data/data_5.txt: the tc.png graph relabelled 0-based
    # (v -> v - 1), so the same 9 tuples, deltas [3, 1, 0], 3 iterations:
    # first join {(0,3) twice, (1,4), (2,4)}, second {(0,4) twice}, third empty.
    r = oracle.tc(read_edges("data_5.txt"))
    assert r["n_out"] == 9
    assert [t[3] for t in r["trace"]] == [3, 1, 0]
    assert r["iterations"] == 3
    assert pairs_of(r["out"]) == {(3, 4), (0, 1), (0, 2), (2, 3), (1, 3), (0, 3), (1, 4),
                                  (2, 4), (0, 4)}


def test_tc_chain_closed_form():
    n = 300
    e = gdl.pack(np.arange(n - 1), np.arange(1, n))
    r = oracle.tc(e)
    assert r["n_out"] == n * (n - 1) // 2
    deltas = [t[3] for t in r["trace"]]
    assert deltas == [n - k - 2 for k in range(n - 1)]
    assert r["iterations"] == n - 1


def distances_closed_form(edges, n_nodes):
    """#pairs (u, v) by shortest walk length >= 1 (cycles for u == v)."""
    s, d = gdl.unpack(edges)
    A = sp.csr_matrix((np.ones(len(s)), (s, d)), shape=(n_nodes, n_nodes))
    D = shortest_path(A, method="D", unweighted=True)
    # shortest cycle through u: 1 + min over in-neighbours p of dist(u, p)
    diag = np.full(n_nodes, np.inf)
    for u, p in zip(d, s):  # edge p -> u
        diag[u] = min(diag[u], (D[u, p] if u != p else 0) + 1)
    np.fill_diagonal(D, diag)
    fin = D[np.isfinite(D)].astype(np.int64)
    return np.bincount(fin)


def test_tc_er_matches_bfs_closed_form():
    # the ER workload (2,000 nodes, 5,000 distinct edges, seed 0)
    e = gdl.gen_er(0, 2000, 5000)
    r = oracle.tc(e)
    hist = distances_closed_form(e, 2000)
    D = len(hist) - 1
    assert r["n_out"] == int(hist.sum())
    assert [t[3] for t in r["trace"]] == [int(hist[k + 2]) for k in range(D - 1)] + [0]
    assert r["iterations"] == D


def test_tc_tree_matches_depth_closed_form():
    e = gdl.gen_tree(7, 3000, 50)
    child, parent = gdl.unpack(e)
    depth = np.zeros(3000, dtype=np.int64)
    for c, p in zip(child, parent):  # parents precede children
        depth[c] = depth[p] + 1
    r = oracle.tc(e)
    D = int(depth.max())
    assert r["n_out"] == int(depth.sum())
    assert [t[3] for t in r["trace"]] == [int((depth >= k + 2).sum()) for k in range(D - 1)] + [0]
    assert r["iterations"] == D


def test_tc_duplicate_rows_seed_delta_from_raw_prefix():
    # copy_relation_container seeds delta with the first |full| raw rows
    # (relation.cu:764-795), so duplicate rows early in the file push unique
    # rows out of the initial delta.  Pure-Python restatement of that rule on a
    # hand-made case: rows [A, A, B], |full| = 2 -> delta0 = {A}.
    A, B, C = (1, 2), (2, 3), (0, 1)
    e = gdl.pack([A[0], A[0], B[0], C[0]], [A[1], A[1], B[1], C[1]])
    r = oracle.tc(e)
    # full = {A, B, C}; delta0 = first 3 raw rows = {A, B}; C never in a delta.
    # body 0: edge ⋈ {A, B}: (0,2) via C+A, (1,3) via A+B -> new {(0,2), (1,3)}
    # body 1: edge ⋈ {(0,2),(1,3)}: (0,3) via C+(1,3) -> new {(0,3)}; body 2: none
    assert [t[3] for t in r["trace"]] == [2, 1, 0]
    assert pairs_of(r["out"]) == {A, B, C, (0, 2), (1, 3), (0, 3)}


# ------------------------------------------------------------------ SG

def sg_matrix_fixpoint(edges, n):
    """S0 = offdiag(A^T A); delta_{k+1} = (A^T delta_k A) minus S (boolean)."""
    s, d = gdl.unpack(edges)
    A = sp.csr_matrix((np.ones(len(s), dtype=np.int64), (s, d)), shape=(n, n))
    At = A.T.tocsr()
    S0 = (At @ A).tolil()
    S0.setdiag(0)
    S = (sp.csr_matrix(S0) > 0).astype(np.int64)
    S.eliminate_zeros()
    delta, deltas = S, []
    while True:
        nxt = ((At @ delta @ A) > 0).astype(np.int64)
        new = nxt - nxt.multiply(S)
        new.eliminate_zeros()
        deltas.append(int(new.nnz))
        if new.nnz == 0:
            break
        S = S + new
        delta = new
    return S, deltas


def test_sg_kat_hipc():
    r = oracle.sg(read_edges("data_5.txt"))
    assert pairs_of(r["out"]) == {(1, 2), (2, 1), (3, 3), (4, 4)}
    assert r["iterations"] == 3


@pytest.mark.parametrize("fname,name", [("data_7035.txt", "OL.cedge"),
                                        ("data_21693.txt", "cal.cedge"),
                                        ("data_23874.txt", "TG.cedge")])
def test_sg_sizes_match_cudf(fname, name):
    sizes = {row[0]: row[2] for row in json.load(open(os.path.join(DATA, "cudf_sg_sizes.json")))}
    r = oracle.sg(read_edges(fname))
    assert r["n_out"] == sizes[name]


def test_sg_matches_matrix_fixpoint():
    e = gdl.gen_rmat(3, 9, 1500)
    r = oracle.sg(e)
    S, deltas = sg_matrix_fixpoint(e, 1 << 9)
    S = S.tocoo()
    assert r["n_out"] == S.nnz
    assert pairs_of(r["out"]) == set(zip(S.row.tolist(), S.col.tolist()))
    assert [t[3] for t in r["trace"]] == deltas


def test_sg_data_7035_matrix_fixpoint_trace():
    e = read_edges("data_7035.txt")
    r = oracle.sg(e)
    n = int(max(gdl.unpack(e)[0].max(), gdl.unpack(e)[1].max())) + 1
    S, deltas = sg_matrix_fixpoint(e, n)
    assert r["n_out"] == S.nnz
    assert [t[3] for t in r["trace"]] == deltas


# ------------------------------------------------------------------ CSPA

def cspa_naive(assign, deref):
    """Naive fixpoint of test/souffle/cspa.dl:25-39 over Python sets."""
    VF = {(y, x) for y, x in assign} | {(x, x) for x, _ in assign} | {(x, x) for _, x in assign}
    MA = {(x, x) for _, x in assign} | {(x, x) for x, _ in assign}
    VA = set()

    def by_first(rel):
        d = {}
        for a, b in rel:
            d.setdefault(a, []).append(b)
        return d

    while True:
        vf1, ma1, va1, as1, de1 = (by_first(VF), by_first(MA), by_first(VA), by_first(assign),
                                   by_first(deref))
        nVF = set(VF)
        nVF |= {(x, y) for x, z in VF for y in vf1.get(z, ())}
        nVF |= {(x, y) for x, zs in as1.items() for z in zs for y in ma1.get(z, ())}
        nVA = set(VA)
        nVA |= {(x, y) for z, xs in vf1.items() for x in xs for y in xs}
        nVA |= {(x, y) for z, xs in vf1.items() for w in ma1.get(z, ()) for x in xs
                for y in vf1.get(w, ())}
        nMA = set(MA)
        nMA |= {(x, w) for y, xs in de1.items() for z in va1.get(y, ()) for x in xs
                for w in de1.get(z, ())}
        if nVF == VF and nVA == VA and nMA == MA:
            return VF, MA, VA
        VF, MA, VA = nVF, nMA, nVA


@pytest.mark.parametrize("seed", range(8))
def test_cspa_matches_naive_datalog(seed):
    rng = np.random.default_rng(seed)
    n_ids = 12 + seed
    assign = {tuple(x) for x in rng.integers(0, n_ids, (10 + 2 * seed, 2)).tolist()}
    deref = {tuple(x) for x in rng.integers(0, n_ids, (8 + seed, 2)).tolist()}
    a = gdl.pack([p[0] for p in assign], [p[1] for p in assign])
    d = gdl.pack([p[0] for p in deref], [p[1] for p in deref])
    r = oracle.cspa(a, d)
    VF, MA, VA = cspa_naive(assign, deref)
    assert pairs_of(r["extra"][0]) == VF
    assert pairs_of(r["extra"][1]) == MA
    assert pairs_of(r["out"]) == VA


@pytest.mark.parametrize("name", ["hpc_talk", "data_7035_dups", "er", "tree", "chain"])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_multicore_cpu_baseline_is_the_same_loop(name, threads):
    # the CPU baseline (oracle/cpu_tc_omp.c) runs the oracle's TC loop split
    # by the path's second column: same size, iterations and per-body delta
    edges = {"hpc_talk": lambda: read_edges("hpc_talk.txt"),
             "data_7035_dups": lambda: read_edges("data_7035.txt"),
             "er": lambda: gdl.gen_er(0, 2000, 5000),
             "tree": lambda: gdl.gen_tree(0, 20000, 1000),
             "chain": lambda: gdl.pack(np.arange(299), np.arange(1, 300))}[name]()
    want = oracle.tc(edges)
    got = oracle.tc_omp(edges, threads)
    assert got["n_out"] == want["n_out"] and got["iterations"] == want["iterations"]
    assert got["deltas"] == [t[3] for t in want["trace"]]
    assert got["derived"] == want["derived"]
    capped = oracle.tc_omp(edges, threads, max_iteration=2)
    assert capped["deltas"] == [t[3] for t in oracle.tc(edges, max_iteration=2)["trace"]]


def test_cspa_delta_conservation():
    a, d = gdl.gen_cspa(5, 150, 1.1, 200, 100)
    r = oracle.cspa(a, d)
    # every update relation's final size = sum of its deltas (all start empty)
    for rel, final in ((0, r["sizes"][0]), (2, r["sizes"][1]), (4, r["sizes"][2])):
        tot = sum(t[3] for t in r["trace_init"] if t[1] == rel) if rel < 4 else 0
        tot += sum(t[3] for t in r["trace"] if t[1] == rel)
        assert tot == final


@pytest.mark.parametrize("name", ["hpc_talk", "data_5", "er", "tree", "chain", "rmat_small"])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_souffle_plan_baselines_give_the_oracle_s_sets(name, threads):
    # oracle/cpu_souffle.c restates Souffle's SG plan and its non-linear TC
    # plan (different loops from the gdlog encodings): same final sets
    edges = {"hpc_talk": lambda: read_edges("hpc_talk.txt"),
             "data_5": lambda: read_edges("data_5.txt"),
             "er": lambda: gdl.gen_er(0, 300, 900),
             "tree": lambda: gdl.gen_tree(0, 3000, 100),
             "chain": lambda: gdl.pack(np.arange(99), np.arange(1, 100)),
             "rmat_small": lambda: gdl.gen_rmat(1, 10, 3000)}[name]()
    tc, sg = oracle.tc(edges), oracle.sg(edges)
    r = oracle.souffle_tc(edges, threads)
    assert r["n_out"] == tc["n_out"] and r["checksum"] == gdl.tuple_checksum(tc["out"])
    assert sum(r["trace"]) == tc["n_out"] - np.unique(edges).size
    s = oracle.souffle_sg(edges, threads)
    assert s["n_out"] == sg["n_out"] and s["checksum"] == gdl.tuple_checksum(sg["out"])
    # the non-recursive rule and the recursive bodies: the same per-body sets as
    # the gdlog encoding's init LIE and sg LIE (both are semi-naive over delta sg)
    want = [sum(t[3] for t in sg["trace_init"])] + [t[3] for t in sg["trace"]]
    strip = lambda xs: xs[:max([i + 1 for i, x in enumerate(xs) if x] or [0])]
    assert strip(s["trace"]) == strip(want)


def test_souffle_tc_plan_is_non_linear():
    # chain of 65 nodes: the linear loop needs 64 bodies, path o path doubles
    # the covered distance each body: 1 + log2(64) bodies derive, one more ends
    e = gdl.pack(np.arange(64), np.arange(1, 65))
    r = oracle.souffle_tc(e, 4)
    assert r["n_out"] == 65 * 64 // 2 and r["iterations"] == 7
    assert oracle.tc(e)["iterations"] == 64


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("threads", [1, 4])
def test_souffle_cspa_plan_gives_the_oracle_s_sets(seed, threads):
    a, d = gdl.gen_cspa(seed, 60 + 40 * seed, 1.1, 90 + 60 * seed, 45 + 30 * seed)
    w = oracle.cspa(a, d)
    r = oracle.souffle_cspa(a, d, threads)
    assert r["sizes"] == [int(w["sizes"][0]), int(w["sizes"][1]), int(w["n_out"])]
    assert r["checksums"] == [gdl.tuple_checksum(w["extra"][0]),
                              gdl.tuple_checksum(w["extra"][1]), gdl.tuple_checksum(w["out"])]
