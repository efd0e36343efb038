"""Program-level parity: the TC / SG / CSPA programs on the gdlog C++ API
(HIP engine, through gdl_programs.h) against the CPU oracle -- bit-exact
sorted result sets, per-iteration delta trace and iteration counts -- and, at
the full workload sizes, against the oracle's golden fixtures."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, gdl, read_edges

pytestmark = pytest.mark.gpu


def sha(keys):
    return hashlib.sha256(np.ascontiguousarray(keys, dtype="<u8").tobytes()).hexdigest()


# auto: FULL as an append log + a membership set (a bitmap when the ids fit
# 16 bits, as on every input here); auto_hash: the same with the hash-table
# set forced (GDLOG_FULL_SET=hash); merge: the reference's sorted FULL
MODES = {"auto": gdl.FULL_MODE_AUTO, "auto_hash": gdl.FULL_MODE_AUTO,
         "merge": gdl.FULL_MODE_MERGE}


@pytest.fixture(params=sorted(MODES))
def mode(request):
    gdl.set_full_mode(MODES[request.param])
    if request.param == "auto_hash":
        os.environ["GDLOG_FULL_SET"] = "hash"
    yield request.param
    os.environ.pop("GDLOG_FULL_SET", None)
    gdl.set_full_mode(gdl.FULL_MODE_AUTO)


def trace_cols(tr, mode):
    # (iteration, relation, newt, delta, full); in auto mode an append-mode
    # relation's newt is not deduplicated before the set difference, so its
    # raw row count is not compared
    if mode == "merge":
        return [(t[0], t[1], t[2], t[3], t[4]) for t in tr]
    return [(t[0], t[1], t[3], t[4]) for t in tr]


def assert_same(got, want, mode="merge"):
    assert got["n_out"] == want["n_out"]
    assert np.array_equal(got["out"], want["out"])
    assert got["iterations"] == want["iterations"]
    assert got["init_iterations"] == want["init_iterations"]
    assert trace_cols(got["trace"], mode) == trace_cols(want["trace"], mode)
    assert trace_cols(got["trace_init"], mode) == trace_cols(want["trace_init"], mode)
    assert got["derived"] == want["derived"]


TC_CASES = {
    "hpc_talk": lambda: read_edges("hpc_talk.txt"),
    "data_5": lambda: read_edges("data_5.txt"),
    "data_7035_dups": lambda: read_edges("data_7035.txt"),
    "chain_300": lambda: gdl.pack(np.arange(299), np.arange(1, 300)),
    "er_500": lambda: gdl.gen_er(3, 500, 1500),
    "tree_5k": lambda: gdl.gen_tree(11, 5000, 100),
    "rmat_12": lambda: gdl.gen_rmat(4, 12, 6000),
    "single_edge": lambda: gdl.pack([1], [2]),
    "self_loop_cycle": lambda: gdl.pack([0, 1, 2, 2], [1, 2, 0, 2]),
}


@pytest.mark.parametrize("name", sorted(TC_CASES))
@pytest.mark.parametrize("on_device", [0, 1])
def test_tc_matches_oracle(ctx, mode, name, on_device):
    e = TC_CASES[name]()
    assert_same(ctx.run_tc(e, on_device=bool(on_device), out_cap=5_000_000), oracle.tc(e), mode)


def test_tc_max_iteration_partial_state(ctx, mode):
    e = gdl.gen_tree(2, 3000, 30)
    for k in (0, 1, 5):
        assert_same(ctx.run_tc(e, max_iteration=k), oracle.tc(e, max_iteration=k), mode)


def test_tc_empty_input(ctx):
    r = ctx.run_tc(np.zeros(0, dtype=np.uint64))
    assert r["n_out"] == 0 and r["iterations"] == 1


SG_CASES = {
    "hipc": lambda: read_edges("data_5.txt"),
    "OL.cedge": lambda: read_edges("data_7035.txt"),
    "cal.cedge": lambda: read_edges("data_21693.txt"),
    "TG.cedge": lambda: read_edges("data_23874.txt"),
    "rmat_10": lambda: gdl.gen_rmat(6, 10, 3000),
}


@pytest.mark.parametrize("name", sorted(SG_CASES))
def test_sg_matches_oracle(ctx, mode, name):
    e = SG_CASES[name]()
    assert_same(ctx.run_sg(e, out_cap=10_000_000), oracle.sg(e), mode)


@pytest.mark.parametrize("fname,size", [("data_48232.txt", 14_503_742),
                                        ("data_26013.txt", 62_056_583)])
def test_sg_large_sizes_match_cudf(ctx, fname, size):
    # result sizes published for these graphs (test/cuDF/sg.json: cti, p2p-Gnutella09)
    r = ctx.run_sg(read_edges(fname), want_keys=False)
    assert r["n_out"] == size
    # every iteration's delta is accounted for in the final size
    assert r["derived"] == size


@pytest.mark.parametrize("seed", range(6))
def test_cspa_matches_oracle(ctx, mode, seed):
    rng = np.random.default_rng(100 + seed)
    n_ids = 20 + 10 * seed
    a = np.unique(gdl.pack(*rng.integers(0, n_ids, (2, 3 * n_ids))))
    rng.shuffle(a)
    d = np.unique(gdl.pack(*rng.integers(0, n_ids, (2, 2 * n_ids))))
    got = ctx.run_cspa(a, d, out_cap=1_000_000, extra_cap=1_000_000)
    want = oracle.cspa(a, d)
    assert_same(got, want, mode)
    assert np.array_equal(got["extra"][0], want["extra"][0])
    assert np.array_equal(got["extra"][1], want["extra"][1])
    assert got["sizes"] == want["sizes"]


def test_cspa_zipf_small_matches_oracle(ctx, mode):
    a, d = gdl.gen_cspa(5, 150, 1.1, 200, 100)
    got = ctx.run_cspa(a, d, out_cap=1_000_000, extra_cap=1_000_000)
    want = oracle.cspa(a, d)
    assert_same(got, want, mode)


def golden():
    p = os.path.join(GOLDEN, "configs.json")
    return json.load(open(p)) if os.path.exists(p) else {}


def workload_input(name):
    if name == "tc_er_2000_5000":
        return gdl.gen_er(0, 2000, 5000)
    if name == "tc_tree_200k":
        return gdl.gen_tree(0, 200000, 1000)
    if name == "sg_rmat15_40k":
        return gdl.gen_rmat(1, 15, 40000)
    if name == "cspa_zipf_500":
        return gdl.gen_cspa(5, 500, 1.1, 750, 375)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["tc_er_2000_5000", "tc_tree_200k", "sg_rmat15_40k",
                                  "cspa_zipf_500"])
def test_full_workload_matches_golden(ctx, mode, name):
    g = golden().get(name)
    if g is None:
        pytest.skip(f"no golden fixture for {name}")
    inp = workload_input(name)
    if g["kind"] == "cspa":
        a, d = inp
        assert sha(np.concatenate([a, d])) == g["input_sha"]
        r = ctx.run_cspa(a, d, on_device=True, out_cap=g["n_out"] + 1,
                         extra_cap=max(g["vf_size"], g["ma_size"]) + 1)
        assert sha(r["extra"][0]) == g["vf_sha"] and sha(r["extra"][1]) == g["ma_sha"]
    else:
        assert sha(inp) == g["input_sha"]
        run = ctx.run_tc if g["kind"] == "tc" else ctx.run_sg
        r = run(inp, on_device=True, out_cap=g["n_out"] + 1)
    assert r["n_out"] == g["n_out"]
    assert sha(r["out"]) == g["out_sha"]
    assert r["iterations"] == g["iterations"]
    assert r["init_iterations"] == g["init_iterations"]
    assert trace_cols(r["trace"], mode) == trace_cols(g["trace"], mode)
    assert trace_cols(r["trace_init"], mode) == trace_cols(g["trace_init"], mode)
    assert r["derived"] == g["derived"]
