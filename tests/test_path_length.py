"""All-pairs hop counts: the reference's path_length program
(test/path_length.cu:49-124, the PLEN executable), whose path relation
carries the hop count as a dependent column.

The oracle (oracle.path_length, a numpy restatement of that loop) is pinned
against breadth-first distances (scipy's unweighted shortest paths; a pair
(x, x) is the shortest cycle through x) on duplicate-free graphs, and its
pairs against the TC oracle on a file with duplicate rows, where the first
delta is the reference's first |FULL| raw rows.  The engine's pairs and hop
counts must equal the oracle's bit for bit."""
import os
import subprocess

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

import oracle
from conftest import gdl, read_edges


def bfs_hops(edges):
    e = np.asarray(edges, dtype=np.uint64)
    s = (e >> np.uint64(32)).astype(np.int64)
    t = (e & np.uint64(0xffffffff)).astype(np.int64)
    n = int(max(s.max(), t.max())) + 1
    g = csr_matrix((np.ones(e.size), (s, t)), shape=(n, n))
    d = shortest_path(g, directed=True, unweighted=True)
    out = {}
    for x in range(n):
        for y in np.nonzero(np.isfinite(d[x]) & (d[x] > 0))[0]:
            out[(x, int(y))] = int(d[x, y])
        # (x, x): an edge x -> a, then back from a to x
        cyc = [1 + d[a, x] for a in t[s == x] if np.isfinite(d[a, x])]
        if cyc:
            out[(x, x)] = int(min(cyc))
    return out


@pytest.mark.parametrize("seed,nodes,m", [(1, 60, 90), (2, 120, 200), (3, 40, 160)])
def test_oracle_hop_counts_are_bfs_distances(seed, nodes, m):
    e = np.unique(gdl.gen_er(seed, nodes, m))
    keys, lens = oracle.path_length(e)
    want = bfs_hops(e)
    got = {(int(k) >> 32, int(k) & 0xffffffff): int(l) for k, l in zip(keys, lens)}
    assert got == want


def test_oracle_pairs_are_the_tc_pairs_with_duplicate_rows():
    e = read_edges("data_7035.txt")
    keys, lens = oracle.path_length(e)
    assert np.array_equal(keys, oracle.tc(e)["out"])
    assert lens.min() == 1 and np.all(np.diff(keys.astype(np.int64)) > 0)


INPUTS = {
    "er_500": lambda: gdl.gen_er(3, 500, 1500),
    "tree_5k": lambda: gdl.gen_tree(11, 5000, 100),
    "data_7035_dups": lambda: read_edges("data_7035.txt"),
    "rmat_10": lambda: gdl.gen_rmat(6, 10, 3000),
    # ids near 2^31 leave no spare bits: the labels take the binary-search path
    "chain_high_ids": lambda: gdl.pack(np.arange(60) + (1 << 31), np.arange(1, 61) + (1 << 31)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(INPUTS))
def test_path_length_matches_the_oracle(ctx, name):
    e = INPUTS[name]()
    got = ctx.run_path_length(e)
    keys, lens = oracle.path_length(e)
    assert np.array_equal(got["out"], keys)
    assert np.array_equal(got["lens"], lens)
    assert got["longest"] == lens.max()
    assert got["iterations"] == oracle.tc(e)["iterations"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(n for n in INPUTS if n != "chain_high_ids"))
@pytest.mark.parametrize("on_device", [0, 1])
def test_reference_encoding_with_hooks_matches_the_oracle(ctx, name, on_device):
    # test/path_length.cu as written: path_3__1_2_3 of arity 3 with one
    # dependent column, the generator hook newt[2] = outer[2] + 1 run on the
    # device, the set difference over (from, to) only
    e = INPUTS[name]()
    got = ctx.run_path_length(e, on_device=bool(on_device), hooks=True)
    keys, lens = oracle.path_length(e)
    assert np.array_equal(got["out"], keys)
    assert np.array_equal(got["lens"], lens)
    assert got["longest"] == lens.max()
    assert got["iterations"] == oracle.tc(e)["iterations"]
    assert got["derived"] == oracle.tc(e)["derived"]


@pytest.mark.gpu
def test_reference_encoding_refuses_ids_its_key_layout_cannot_hold(ctx):
    # from and to take 32 bits each beside the hop count: nothing is left for it
    with pytest.raises(gdl.GdlError) as err:
        ctx.run_path_length(INPUTS["chain_high_ids"](), hooks=True)
    assert err.value.status == 4  # unsupported, before anything runs
    small = ctx.run_path_length(gdl.pack([4], [4]), hooks=True)
    assert small["out"].tolist() == [(4 << 32) | 4] and small["lens"].tolist() == [1]


@pytest.mark.gpu
def test_path_length_edge_cases(ctx):
    empty = ctx.run_path_length(np.zeros(0, dtype=np.uint64))
    assert empty["n_out"] == 0
    one = ctx.run_path_length(gdl.pack(np.array([4]), np.array([4])))  # a self loop
    assert one["out"].tolist() == [(4 << 32) | 4] and one["lens"].tolist() == [1]
    with pytest.raises(gdl.GdlError) as err:
        ctx.run_path_length(gdl.gen_er(3, 500, 1500), out_cap=10)
    assert err.value.status == 5  # overflow


@pytest.mark.gpu
def test_plen_executable_writes_hop_counts(tmp_path):
    e = gdl.gen_tree(2, 2000, 30)
    fn = tmp_path / "g.tsv"
    with open(fn, "w") as f:
        for k in e:
            f.write(f"{int(k) >> 32}\t{int(k) & 0xffffffff}\n")
    exe = os.path.join(os.path.dirname(gdl.__file__), "lib", "PLEN")
    r = subprocess.run([exe, str(fn), "--out", str(tmp_path / "o.tsv")], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = np.loadtxt(tmp_path / "o.tsv", dtype=np.uint64, ndmin=2)
    keys, lens = oracle.path_length(e)
    assert np.array_equal(gdl.pack(rows[:, 0], rows[:, 1]), keys)
    assert np.array_equal(rows[:, 2].astype(np.uint32), lens)
    assert f"longest path {lens.max()}" in r.stdout
