"""Writes tests/golden/cspa_stated.json: CSPA at the stated input ratio (3
assign and 1.5 dereference facts per id, Zipf 1.1 sources, uniform targets,
seed 5) on 1,000 ids -- the largest such input whose GPU run completes
within a minute -- from oracle/cpu_souffle.c's restatement of Souffle's CSPA
plan (the gdlog-encoding oracle's raw joins need more host memory than a
machine has at this size; the restatement is pinned against that oracle on
small inputs in tests/test_oracle.py).

    python tests/golden/make_cspa_golden.py [--ids 1000] [--threads 8]
"""
import argparse
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ids", type=int, default=1000)
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    from conftest import gdl
    import oracle
    n = args.ids
    a, d = gdl.gen_cspa(5, n, 1.1, 3 * n, 3 * n // 2)
    t = time.time()
    r = oracle.souffle_cspa(a, d, args.threads)
    rec = dict(ids=n, seed=5, zipf=1.1, n_assign=int(a.size), n_deref=int(d.size),
               input_sha=hashlib.sha256(np.concatenate([a, d]).astype("<u8").tobytes())
               .hexdigest(),
               value_flow=r["sizes"][0], memory_alias=r["sizes"][1], value_alias=r["sizes"][2],
               checksums=[str(c) for c in r["checksums"]], recursive_bodies=r["iterations"],
               seconds=round(time.time() - t, 1))
    with open(os.path.join(HERE, "cspa_stated.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print(rec)


if __name__ == "__main__":
    main()
