"""The TC / SG / CSPA executables (the reference's drivers of the same names):
tab-separated input files in, sorted tab-separated and binary results out,
compared with the oracle on the same files; the printed checksum is the
order-independent one of the result; --checkpoint with --max-iteration, then
--resume, finishes the same fixpoint."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import DATA, gdl, read_edges

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(gdl.__file__), "lib")


def run(prog, *args):
    r = subprocess.run([os.path.join(BIN, prog), *map(str, args)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def read_tsv(path):
    rows = np.loadtxt(path, dtype=np.uint64, ndmin=2, delimiter="\t")
    if rows.size == 0:
        return np.zeros(0, dtype=np.uint64)
    return gdl.pack(rows[:, 0], rows[:, 1])


def checksum_line(stdout):
    line = [x for x in stdout.splitlines() if x.startswith("checksum ")][-1]
    return int(line.split()[1], 16)


def write_pairs(path, keys):
    with open(path, "w") as f:
        for k in keys:
            f.write(f"{int(k) >> 32}\t{int(k) & 0xffffffff}\n")


@pytest.mark.parametrize("prog", ["TC", "SG"])
def test_tc_sg_outputs_match_the_oracle(tmp_path, prog):
    fn = os.path.join(DATA, "data_7035.txt")  # with duplicate rows: the raw-row quirk
    edges = read_edges("data_7035.txt")
    want = (oracle.tc if prog == "TC" else oracle.sg)(edges)["out"]
    out = run(prog, fn, "--out", tmp_path / "o.tsv", "--out-bin", tmp_path / "o.bin")
    assert np.array_equal(read_tsv(tmp_path / "o.tsv"), want)
    assert np.array_equal(np.fromfile(tmp_path / "o.bin", dtype="<u8"), want)
    assert checksum_line(out) == gdl.tuple_checksum(want)


def test_cspa_writes_its_three_relations(tmp_path):
    a, d = gdl.gen_cspa(5, 150, 1.1, 200, 100)
    write_pairs(tmp_path / "assign.facts", a)
    write_pairs(tmp_path / "dereference.facts", d)
    res = tmp_path / "out"
    res.mkdir()
    out = run("CSPA", tmp_path, "--out", res)
    want = oracle.cspa(a, d)
    assert np.array_equal(read_tsv(res / "ValueFlow.csv"), want["extra"][0])
    assert np.array_equal(read_tsv(res / "MemoryAlias.csv"), want["extra"][1])
    assert np.array_equal(read_tsv(res / "ValueAlias.csv"), want["out"])
    assert checksum_line(out) == gdl.tuple_checksum(want["out"])


def test_tc_checkpoint_then_resume(tmp_path):
    fn = tmp_path / "tree.tsv"
    edges = gdl.gen_tree(11, 3000, 60)
    write_pairs(fn, edges)
    want = oracle.tc(edges)
    ck = tmp_path / "ck"
    ck.mkdir()
    part = run("TC", fn, "--checkpoint", ck, "--max-iteration", 5)
    assert "iterations 6 " in part
    out = run("TC", fn, "--checkpoint", ck, "--resume", "--out", tmp_path / "o.tsv")
    assert np.array_equal(read_tsv(tmp_path / "o.tsv"), want["out"])
    assert f"iterations {want['iterations']} " in out
