"""Writes tests/golden/configs.json: the CPU oracle's results on the
workloads, so GPU tests at full size compare against fixed numbers instead of
re-running the oracle.

    python tests/golden/make_golden.py [name ...]

Per workload: input sha256 (packed u64 keys, generation order), result size,
sha256 of the sorted result keys, iteration counts, per-iteration delta trace,
derived tuples.  The generators are the package's SplitMix64 ones.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

from conftest import gdl  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(HERE, "configs.json")


def sha(keys):
    return hashlib.sha256(np.ascontiguousarray(keys, dtype="<u8").tobytes()).hexdigest()


def workloads():
    return {
        # 2,000-node directed Erdos-Renyi graph, 5,000 distinct edges, seed 0
        "tc_er_2000_5000": ("tc", lambda: gdl.gen_er(0, 2000, 5000)),
        # 200,000-node child->parent tree, parent window 1,000, seed 0
        "tc_tree_200k": ("tc", lambda: gdl.gen_tree(0, 200000, 1000)),
        # R-MAT scale 15, 40,000 distinct edges, 0.57/0.19/0.19/0.05, seed 1
        "sg_rmat15_40k": ("sg", lambda: gdl.gen_rmat(1, 15, 40000)),
        # CSPA at a reduced scale the full fixpoint fits (see the CSPA notes)
        "cspa_zipf_500": ("cspa", lambda: gdl.gen_cspa(5, 500, 1.1, 750, 375)),
    }


def run(name, kind, make):
    inp = make()
    t = time.time()
    if kind == "tc":
        r = oracle.tc(inp)
        rec = dict(kind=kind, n_edges=int(inp.size), input_sha=sha(inp))
    elif kind == "sg":
        r = oracle.sg(inp)
        rec = dict(kind=kind, n_edges=int(inp.size), input_sha=sha(inp))
    else:
        a, d = inp
        r = oracle.cspa(a, d)
        rec = dict(kind=kind, n_assign=int(a.size), n_deref=int(d.size),
                   input_sha=sha(np.concatenate([a, d])),
                   vf_size=int(r["sizes"][0]), ma_size=int(r["sizes"][1]),
                   vf_sha=sha(r["extra"][0]), ma_sha=sha(r["extra"][1]))
    rec.update(n_out=int(r["n_out"]), out_sha=sha(r["out"]), iterations=int(r["iterations"]),
               init_iterations=int(r["init_iterations"]), derived=int(r["derived"]),
               trace=[list(map(int, x)) for x in r["trace"]],
               trace_init=[list(map(int, x)) for x in r["trace_init"]],
               oracle_seconds=round(time.time() - t, 2))
    return rec


def main(names):
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name, (kind, make) in workloads().items():
        if names and name not in names:
            continue
        rec = run(name, kind, make)
        data[name] = rec
        print(name, rec["n_out"], rec["iterations"], rec["oracle_seconds"], "s", flush=True)
        json.dump(data, open(OUT, "w"), indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
