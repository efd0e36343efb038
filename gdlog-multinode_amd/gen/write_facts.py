"""Writes the seeded synthetic inputs as tab-separated fact files, for the
command-line drivers (TC, SG, PLEN read one edge file; CSPA and DL read a
directory of <relation>.facts).

    python gdlog-multinode_amd/gen/write_facts.py tree DIR [--nodes 200000] [--window 1000]
    python gdlog-multinode_amd/gen/write_facts.py er   DIR [--nodes 2000] [--edges 5000]
    python gdlog-multinode_amd/gen/write_facts.py rmat DIR [--scale 15] [--edges 40000]
    python gdlog-multinode_amd/gen/write_facts.py cspa DIR [--ids 500] [--assign 750]
                                                          [--deref 375] [--zipf 1.1]

Graphs go to DIR/edge.facts, CSPA inputs to DIR/assign.facts and
DIR/dereference.facts; --seed picks the SplitMix64 seed (default 0 for the
graphs, 5 for CSPA: the benchmark workloads).
"""
import argparse
import importlib.util
import os
import sys

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _pkg():
    spec = importlib.util.spec_from_file_location("gdlog_amd", os.path.join(PKG, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gdlog_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def write(path, keys):
    k = np.asarray(keys, dtype=np.uint64)
    rows = np.stack([k >> np.uint64(32), k & np.uint64(0xffffffff)], axis=1)
    np.savetxt(path, rows, fmt="%d", delimiter="\t")
    print(f"{path}: {k.size} rows")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("kind", choices=["tree", "er", "rmat", "cspa"])
    ap.add_argument("dir")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--nodes", type=int, default=None)
    ap.add_argument("--window", type=int, default=1000)
    ap.add_argument("--edges", type=int, default=None)
    ap.add_argument("--scale", type=int, default=15)
    ap.add_argument("--ids", type=int, default=500)
    ap.add_argument("--assign", type=int, default=750)
    ap.add_argument("--deref", type=int, default=375)
    ap.add_argument("--zipf", type=float, default=1.1)
    a = ap.parse_args()
    gdl = _pkg()
    os.makedirs(a.dir, exist_ok=True)
    if a.kind == "cspa":
        seed = 5 if a.seed is None else a.seed
        assign, deref = gdl.gen_cspa(seed, a.ids, a.zipf, a.assign, a.deref)
        write(os.path.join(a.dir, "assign.facts"), assign)
        write(os.path.join(a.dir, "dereference.facts"), deref)
        return
    seed = 0 if a.seed is None else a.seed
    if a.kind == "tree":
        edges = gdl.gen_tree(seed, a.nodes or 200_000, a.window)
    elif a.kind == "er":
        edges = gdl.gen_er(seed, a.nodes or 2000, a.edges or 5000)
    else:
        edges = gdl.gen_rmat(1 if a.seed is None else a.seed, a.scale, a.edges or 40000)
    write(os.path.join(a.dir, "edge.facts"), edges)


if __name__ == "__main__":
    main()
