"""The TC / SG / CSPA executables (the reference's drivers of the same names):
tab-separated input files in, sorted tab-separated and binary results out,
compared with the oracle on the same files; the printed checksum is the
order-independent one of the result; --checkpoint with --max-iteration, then
--resume, finishes the same fixpoint.  DL, the Datalog front end's driver,
runs the package's tc.dl / sg.dl / cspa.dl on .facts files to the same
results."""
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


# ---- DL: the Datalog front end's driver ----------------------------------------

DL_DIR = os.path.join(os.path.dirname(gdl.__file__), "programs", "dl")


def relation_lines(stdout):
    """{relation: (size, checksum)} of DL's report lines."""
    out = {}
    for line in stdout.splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[2] == "checksum":
            out[parts[0]] = (int(parts[1]), int(parts[3], 16))
    return out


@pytest.mark.parametrize("prog", ["tc", "sg", "cspa"])
def test_dl_programs_match_the_oracle(tmp_path, prog):
    facts, res = tmp_path / "facts", tmp_path / "out"
    facts.mkdir()
    res.mkdir()
    if prog == "cspa":
        a, d = gdl.gen_cspa(5, 150, 1.1, 200, 100)
        a, d = np.unique(a), np.unique(d)
        write_pairs(facts / "assign.facts", a)
        write_pairs(facts / "dereference.facts", d)
        o = oracle.cspa(a, d)
        want = {"ValueFlow": o["extra"][0], "MemoryAlias": o["extra"][1], "ValueAlias": o["out"]}
    else:
        e = np.unique(gdl.gen_er(3, 500, 1500) if prog == "tc" else gdl.gen_rmat(6, 10, 3000))
        write_pairs(facts / "edge.facts", e)
        want = {"path" if prog == "tc" else "sg": (oracle.tc if prog == "tc" else oracle.sg)(e)["out"]}
    out = run("DL", os.path.join(DL_DIR, prog + ".dl"), "-F", facts, "-D", res, "--plan")
    assert out.startswith("lie ")
    lines = relation_lines(out)
    assert sorted(lines) == sorted(want)
    for rel, keys in want.items():
        assert np.array_equal(read_tsv(res / f"{rel}.csv"), keys), rel
        assert lines[rel] == (keys.size, gdl.tuple_checksum(keys)), rel


def test_dl_checkpoint_then_resume_restores_the_result(tmp_path):
    facts, ck = tmp_path / "facts", tmp_path / "ck"
    facts.mkdir()
    ck.mkdir()
    e = gdl.gen_tree(11, 3000, 60)
    write_pairs(facts / "edge.facts", e)
    want = oracle.tc(e)["out"]
    first = run("DL", os.path.join(DL_DIR, "tc.dl"), "-F", facts, "--checkpoint", ck)
    assert sorted(os.listdir(ck)) == ["lie0.ckpt", "lie1.ckpt"]
    again = run("DL", os.path.join(DL_DIR, "tc.dl"), "-F", facts, "--checkpoint", ck, "--resume")
    assert relation_lines(first) == relation_lines(again) == {
        "path": (want.size, gdl.tuple_checksum(want))}


def test_dl_reports_compile_errors(tmp_path):
    bad = tmp_path / "bad.dl"
    bad.write_text(".decl r(a: number, b: number)\nr(x, y) :- r(x, y), !r(y, x).\n")
    r = subprocess.run([os.path.join(BIN, "DL"), str(bad)], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 1 and "line 2: negation is not supported" in r.stderr


def test_dl_unary_facts_and_results(tmp_path):
    prog = tmp_path / "reach.dl"
    prog.write_text(".decl edge(a: number, b: number)\n.input edge\n.decl start(a: number)\n"
                    ".input start\n.decl reach(a: number)\n.output reach\n"
                    "reach(x) :- start(x).\nreach(y) :- reach(x), edge(x, y).\n")
    (tmp_path / "edge.facts").write_text("1\t2\n2\t3\n3\t1\n7\t8\n")
    (tmp_path / "start.facts").write_text("2\n")
    out = run("DL", prog, "-F", tmp_path, "-D", tmp_path)
    assert (tmp_path / "reach.csv").read_text() == "1\n2\n3\n"
    keys = np.array([1 << 32, 2 << 32, 3 << 32], dtype=np.uint64)
    assert relation_lines(out) == {"reach": (3, gdl.tuple_checksum(keys))}


def test_dl_runs_a_slog_plan(tmp_path):
    from test_datalog import TC_SLOGC
    plan = tmp_path / "tc.slogc"
    plan.write_text(TC_SLOGC)
    e = np.unique(gdl.gen_er(3, 500, 1500))
    write_pairs(tmp_path / "edge.facts", e)
    out = run("DL", plan, "-F", tmp_path, "-D", tmp_path, "--plan")
    assert "path(v1, v2) :- path(v1, v0), edge(v0, v2)." in out
    want = oracle.tc(e)["out"]
    assert np.array_equal(read_tsv(tmp_path / "path.csv"), want)
    assert relation_lines(out) == {"path": (want.size, gdl.tuple_checksum(want))}


def test_dl_cspa_checkpoint_then_resume(tmp_path):
    # a plan with mirrors and tmp relations: each LIE's snapshot restores it
    facts, ck = tmp_path / "facts", tmp_path / "ck"
    facts.mkdir()
    ck.mkdir()
    a, d = gdl.gen_cspa(5, 150, 1.1, 200, 100)
    write_pairs(facts / "assign.facts", a)
    write_pairs(facts / "dereference.facts", d)
    prog = os.path.join(DL_DIR, "cspa.dl")
    first = run("DL", prog, "-F", facts, "--checkpoint", ck, "--every", 2)
    assert sorted(os.listdir(ck)) == ["lie0.ckpt", "lie1.ckpt"]
    again = run("DL", prog, "-F", facts, "--checkpoint", ck, "--resume")
    want = oracle.cspa(a, d)
    lines = relation_lines(again)
    assert lines == relation_lines(first)
    assert lines["ValueAlias"] == (want["out"].size, gdl.tuple_checksum(want["out"]))


def test_dl_reads_the_input_filename_attribute(tmp_path):
    # Souffle's .input filename attribute: relative to -F unless absolute
    e = np.unique(gdl.gen_er(4, 300, 700))
    (tmp_path / "sub").mkdir()
    write_pairs(tmp_path / "sub" / "graph.tsv", e)
    want = oracle.tc(e)["out"]
    for name in ["sub/graph.tsv", str(tmp_path / "sub" / "graph.tsv")]:
        prog = tmp_path / "tc.dl"
        prog.write_text(".decl edge(a: number, b: number)\n"
                        f'.input edge(IO=file, filename="{name}", deliminator="\\t")\n'
                        ".decl path(a: number, b: number)\n.printsize path\n"
                        "path(x, y) :- edge(x, y).\npath(x, y) :- path(x, z), edge(z, y).\n")
        out = run("DL", prog, "-F", tmp_path)
        assert relation_lines(out) == {"path": (want.size, gdl.tuple_checksum(want))}, name
