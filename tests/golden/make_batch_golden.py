"""Writes tests/golden/batch256.json: the CPU oracle's result for every
instance of the 256-instance batched-TC workload.

    python tests/golden/make_batch_golden.py [--workers 4]

Instance k (k = 0..255) is the batched-TC generator at seed k: edge count
log-uniform over 1,000-50,000, an Erdos-Renyi digraph (nodes = edges / 2.5)
or a child->parent tree (parent window 1,000), 50/50.  Per instance: node and
edge counts, input sha256 (packed u64 keys in generation order), |path|,
iteration count, derived tuples, the per-body (|delta|, |full|) trace, the
order-independent checksum (sum of fmix64 of the keys, mod 2^64) and the
sha256 of the sorted path keys.  Every number comes from oracle.tc, the CPU
restatement of the reference's semi-naive loop; the instances run in a
process pool because the oracle is single-threaded.
"""
import argparse
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

OUT = os.path.join(HERE, "batch256.json")
N_INST, MIN_EDGES, MAX_EDGES, WINDOW = 256, 1000, 50000, 1000


def sha(keys):
    return hashlib.sha256(np.ascontiguousarray(keys, dtype="<u8").tobytes()).hexdigest()


def one(q):
    from conftest import gdl
    import oracle
    is_tree, m, nodes = gdl.gen_batched_shape(q, MIN_EDGES, MAX_EDGES)
    e = gdl.gen_batched_instance(q, MIN_EDGES, MAX_EDGES, WINDOW)
    t = time.time()
    r = oracle.tc(e)
    rec = dict(qid=q, tree=bool(is_tree), nodes=int(nodes), n_edges=int(e.size),
               input_sha=sha(e), n_out=int(r["n_out"]), iterations=int(r["iterations"]),
               derived=int(r["derived"]),
               trace=[[int(x[3]), int(x[4])] for x in r["trace"]],
               checksum=str(gdl.tuple_checksum(r["out"])), out_sha=sha(r["out"]),
               oracle_seconds=round(time.time() - t, 2))
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=4)
    args = ap.parse_args()
    from conftest import gdl
    # largest first, so the pool's tail is short
    order = sorted(range(N_INST), key=lambda q: -gdl.gen_batched_shape(q, MIN_EDGES,
                                                                         MAX_EDGES)[1])
    recs = {}
    t0 = time.time()
    with ProcessPoolExecutor(args.workers) as pool:
        for rec in pool.map(one, order):
            recs[rec["qid"]] = rec
            print(f"{len(recs)}/{N_INST} q={rec['qid']} n_out={rec['n_out']} "
                  f"it={rec['iterations']} {rec['oracle_seconds']} s", flush=True)
    data = dict(workload=f"tc_batched_{N_INST}", min_edges=MIN_EDGES, max_edges=MAX_EDGES,
                tree_window=WINDOW, seeds="0..255 (instance k has seed k)",
                instances=[recs[q] for q in range(N_INST)],
                total_derived=sum(r["derived"] for r in recs.values()),
                total_out=sum(r["n_out"] for r in recs.values()),
                oracle_wall_seconds=round(time.time() - t0, 1))
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0)
    print("wrote", OUT, data["total_derived"], "derived", data["oracle_wall_seconds"], "s")


if __name__ == "__main__":
    main()
