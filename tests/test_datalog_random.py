"""Randomly generated Datalog programs through the front end, against the
independent evaluator (tests/datalog_naive.py) on small graphs.

Each program has two binary inputs, a unary input and three or four derived
relations (unary or binary), with rules of one to four atoms chained by
shared variables, drawn with recursion, mutual recursion, reversed reads,
constants and comparisons.  Programs outside the front end's subset must be
refused with GDL_ERR_UNSUPPORTED (never crash); the others must give the
evaluator's relations exactly.  The seeds are fixed, so the set of programs
is the same on every run.
"""
import numpy as np
import pytest

from conftest import gdl
from datalog_naive import evaluate

OPS = ["=", "!=", "<", "<=", ">", ">="]
N_PROGRAMS = 150


def random_program(seed):
    rng = np.random.default_rng(seed)
    arity = {"e": 2, "f": 2, "s": 1}
    derived = [f"d{i}" for i in range(int(rng.integers(3, 5)))]
    for d in derived:
        arity[d] = int(rng.choice([1, 2], p=[0.3, 0.7]))
    rels = list(arity)
    rules = []
    for d in derived:
        for _ in range(int(rng.integers(1, 4))):
            k = int(rng.integers(1, 5))
            nv = 0
            body, vars_ = [], []

            def fresh():
                nonlocal nv
                nv += 1
                return f"v{nv}"
            prev = None
            for _ in range(k):
                r = str(rng.choice(rels))
                if arity[r] == 1:
                    v = prev if prev is not None else fresh()
                    body.append((r, [v]))
                    vars_.append(v)
                    prev = v
                    continue
                shared = prev if prev is not None else fresh()
                other = fresh()
                args = [shared, other] if rng.random() < 0.5 else [other, shared]
                if prev is not None and rng.random() < 0.15:  # a constant instead of `other`
                    args = [str(int(rng.integers(0, 8))) if a == other else a for a in args]
                else:
                    vars_.append(other)
                body.append((r, args))
                vars_.append(shared)
                prev = other if other in vars_ else shared
            pool = sorted(set(vars_), key=lambda v: int(v[1:]))
            head = [str(rng.choice(pool)) for _ in range(arity[d])]
            text = f"{d}({', '.join(head)}) :- " + ", ".join(
                f"{r}({', '.join(a)})" for r, a in body)
            if rng.random() < 0.25:
                lhs = head[0]
                rhs = head[-1] if rng.random() < 0.5 else str(int(rng.integers(0, 12)))
                text += f", {lhs} {rng.choice(OPS)} {rhs}"
            rules.append(text + ".")
    decls = "".join(f".decl {r}({', '.join(f'c{i}: number' for i in range(arity[r]))})\n"
                    for r in rels)
    src = decls + ".input e, f, s\n" + "\n".join(rules) + "\n"
    return src


def random_inputs(seed):
    rng = np.random.default_rng(1000 + seed)
    e = np.unique(gdl.gen_er(seed, 14, 40))
    f = np.unique(gdl.gen_er(seed + 7, 14, 18))
    s = np.unique(rng.integers(0, 14, size=4).astype(np.uint64) << np.uint64(32))
    return dict(e=e, f=f, s=s)


def compiled_programs():
    out = []
    for seed in range(N_PROGRAMS):
        src = random_program(seed)
        try:
            gdl.Datalog(src)
        except gdl.GdlError as err:
            assert err.status == 4, (seed, str(err), src)  # outside the subset, by status
            continue
        out.append(seed)
    return out


def test_random_programs_compile_or_are_refused():
    ok = compiled_programs()
    assert len(ok) >= N_PROGRAMS // 2, len(ok)


@pytest.mark.gpu
def test_random_programs_match_the_naive_evaluator(ctx):
    ran = 0
    for seed in compiled_programs():
        src = random_program(seed)
        inputs = random_inputs(seed)
        want = evaluate(src, inputs)
        got = gdl.Datalog(src).run(ctx, inputs)
        assert sorted(got) == sorted(want), seed
        for rel in want:
            assert np.array_equal(got[rel], want[rel]), (seed, rel, src)
        ran += 1
    assert ran >= N_PROGRAMS // 2
