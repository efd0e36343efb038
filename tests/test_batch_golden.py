"""The 256-instance batched-TC workload at its stated size (edge counts
log-uniform over 1,000-50,000, ER digraphs or child->parent trees, seeds
0..255) through one continuous TcBatch, every instance checked against the
CPU oracle's result recorded in tests/golden/batch256.json
(tests/golden/make_batch_golden.py): |path|, iteration count, derived tuples,
the per-body (|delta|, |full|) trace and the order-independent checksum of the
path set; for the smaller instances also the sha256 of the sorted path keys.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import gdl

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch256.json")


def _golden():
    with open(GOLD) as f:
        return json.load(f)


def _instances(g):
    out = []
    for rec in g["instances"]:
        q = rec["qid"]
        e = gdl.gen_batched_instance(q, g["min_edges"], g["max_edges"], g["tree_window"])
        out.append((q, e, rec["nodes"]))
    return out


def _sha(keys):
    return hashlib.sha256(np.ascontiguousarray(keys, dtype="<u8").tobytes()).hexdigest()


def test_golden_inputs_are_the_generator_s():
    g = _golden()
    assert len(g["instances"]) == 256
    for (q, e, n), rec in zip(_instances(g), g["instances"]):
        is_tree, m, nodes = gdl.gen_batched_shape(q, g["min_edges"], g["max_edges"])
        assert (bool(is_tree), nodes) == (rec["tree"], rec["nodes"])
        assert e.size == rec["n_edges"] and _sha(e) == rec["input_sha"], q
        # conservation: the derived tuples are the deltas of the trace
        assert sum(d for d, _ in rec["trace"]) == rec["derived"] == rec["n_out"] - e.size
        assert rec["trace"][-1][1] == rec["n_out"] and rec["trace"][-1][0] == 0
        assert len(rec["trace"]) == rec["iterations"]


@pytest.mark.gpu
def test_batch_of_256_matches_the_oracle(ctx):
    g = _golden()
    inst = _instances(g)
    with gdl.TcBatch(ctx, keep_keys=False) as b:
        done = b.run(inst)
    bad = []
    for rec in g["instances"]:
        r = done[rec["qid"]]
        got = (r["n_out"], r["iterations"], r["derived"], str(r["checksum"]),
               [[int(t[3]), int(t[4])] for t in r["trace"]])
        want = (rec["n_out"], rec["iterations"], rec["derived"], rec["checksum"], rec["trace"])
        if got != want:
            bad.append(rec["qid"])
    assert not bad, f"instances differing from the oracle: {bad}"


@pytest.mark.gpu
def test_batch_sorted_keys_match_the_oracle_sha(ctx):
    # the 96 instances with the fewest path tuples, their keys returned sorted
    g = _golden()
    small = sorted(g["instances"], key=lambda r: r["n_out"])[:96]
    want = {r["qid"]: r for r in small}
    inst = [x for x in _instances(g) if x[0] in want]
    with gdl.TcBatch(ctx, keep_keys=True) as b:
        done = b.run(inst)
    for q, rec in want.items():
        keys = done[q]["keys"]
        assert keys.size == rec["n_out"], q
        assert _sha(keys) == rec["out_sha"], q
