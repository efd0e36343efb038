"""A pure-Python evaluator of the Datalog subset the engine's front end takes
(TEST INFRASTRUCTURE: the checker of gdl.Datalog on small inputs).

Written independently of the front end: regex parsing, then global
semi-naive evaluation over Python sets (every rule, every body position read
from the previous round's delta, the other atoms from the full relations),
which for positive Datalog reaches the least fixpoint without stratification.
Its results are pinned against the C oracle's TC, SG and CSPA (the restated
reference programs) in test_datalog.py.
"""
import re

import numpy as np

_ATOM = r"([A-Za-z_?][\w?]*)\s*\(\s*([\w?]+)\s*(?:,\s*([\w?]+)\s*)?\)"
_CMP = {"=": lambda a, b: a == b, "!=": lambda a, b: a != b, "<": lambda a, b: a < b,
        "<=": lambda a, b: a <= b, ">": lambda a, b: a > b, ">=": lambda a, b: a >= b}


def parse(src):
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    decls = re.findall(r"\.decl\s+([\w?]+)", src)
    inputs = set()
    kept = []
    for kind, names in re.findall(r"\.(input|output|printsize)\s+([^\n]*)", src):
        for n in re.findall(r"([\w?]+)\s*(?:\([^)]*\))?", names):
            (inputs.add(n) if kind == "input" else kept.append(n))
    text = re.sub(r"\.(decl|input|output|printsize)[^\n]*", " ", src)
    rules, facts = [], []
    for stmt in re.split(r"\.\s", text + " "):
        if ":-" not in stmt:
            f = re.fullmatch(r"\s*" + _ATOM + r"\s*", stmt)
            if f:  # a ground atom
                facts.append((f.group(1), tuple(int(c) for c in f.groups()[1:] if c)))
            continue
        head, body = stmt.split(":-", 1)
        h = re.fullmatch(r"\s*" + _ATOM + r"\s*", head)
        atoms = [(a, tuple(v for v in (x, y) if v)) for a, x, y in re.findall(_ATOM, body)]
        rest = re.sub(_ATOM, " ", body)
        cmps = re.findall(r"([\w?]+)\s*(!=|<=|>=|=|<|>)\s*([\w?]+)", rest)
        hv = tuple(v for v in (h.group(2), h.group(3)) if v)
        rules.append(((h.group(1), hv), atoms, cmps))
    return decls, inputs, kept, rules, facts


def _matches(env, atom, tuples):
    _, args = atom
    for t in tuples:
        e = dict(env)
        ok = True
        for v, val in zip(args, t):
            if v == "_":
                continue
            if v.isdigit():  # a constant
                if int(v) != val:
                    ok = False
                    break
                continue
            if v in e and e[v] != val:
                ok = False
                break
            e[v] = val
        if ok:
            yield e


def _fire(rule, sources):
    (_, hv), atoms, cmps = rule
    envs = [{}]
    for atom, tuples in zip(atoms, sources):
        envs = [e2 for e in envs for e2 in _matches(e, atom, tuples)]
        if not envs:
            return set()
    out = set()
    for e in envs:
        val = lambda v: int(v) if v.isdigit() else e[v]  # noqa: E731 (an id or a variable)
        if all(_CMP[op](val(l), val(r)) for l, op, r in cmps):
            out.add(tuple(e[v] for v in hv))
    return out


def evaluate(src, inputs):
    """inputs: {relation: packed uint64 keys}.  Returns {kept relation: sorted
    packed keys} (every non-input relation when none is kept).  A unary
    relation's tuples are the keys c0 << 32."""
    decls, in_names, kept, rules, facts = parse(src)
    arity = {n: len(a) for n, a in re.findall(r"\.decl\s+([\w?]+)\s*\(([^)]*)\)", src)
             for a in [a.split(",")]}
    full = {n: set() for n in decls}
    for n, keys in inputs.items():
        k = np.asarray(keys, dtype=np.uint64)
        hi, lo = (k >> np.uint64(32)).tolist(), (k & np.uint64(0xffffffff)).tolist()
        full[n] = set(zip(hi, lo)) if arity[n] == 2 else {(x,) for x in hi}
    for n, t in facts:
        full[n].add(t)
    delta = {n: set(s) for n, s in full.items()}
    first = True
    while True:
        new = {n: set() for n in decls}
        for rule in rules:
            head = rule[0][0]
            atoms = rule[1]
            if first:
                new[head] |= _fire(rule, [full[a] for a, _ in atoms])
                continue
            for i, (a, _) in enumerate(atoms):
                if not delta[a]:
                    continue
                srcs = [delta[b] if j == i else full[b] for j, (b, _) in enumerate(atoms)]
                new[head] |= _fire(rule, srcs)
        first = False
        delta = {n: new[n] - full[n] for n in decls}
        if not any(delta.values()):
            break
        for n in decls:
            full[n] |= delta[n]
    names = kept or [n for n in decls if n not in in_names]
    out = {}
    for n in names:
        t = sorted(full[n])
        out[n] = np.array([(x[0] << 32) | (x[1] if len(x) > 1 else 0) for x in t],
                          dtype=np.uint64)
    return out
