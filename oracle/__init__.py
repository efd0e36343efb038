"""CPU oracle (TEST INFRASTRUCTURE ONLY).

ctypes wrapper of oracle/lib/libgdl_oracle.so, the plain-C restatement of the
reference engine's semi-naive algorithm (see gdl_oracle.h for what each
function restates and how it is pinned).  path_length() is a numpy
restatement of test/path_length.cu's loop; tests/test_path_length.py pins it
against breadth-first distances (scipy) and its pairs against tc().  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may use it, and only as the checker / the CPU
baseline -- never as the product path.
"""
import ctypes
import os
from ctypes import POINTER, Structure, c_int, c_int32, c_size_t, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libgdl_oracle.so")

_U64P = POINTER(c_uint64)


class TraceRec(Structure):
    _fields_ = [("iteration", c_int), ("rel", c_int), ("newt", c_uint64), ("delta", c_uint64),
                ("full", c_uint64)]


class OpRec(Structure):
    _fields_ = [("iteration", c_int), ("op", c_int), ("ran", c_int), ("raw", c_uint64),
                ("out_newt", c_uint64)]


class Result(Structure):
    _fields_ = [("iterations", c_int), ("init_iterations", c_int), ("derived", c_uint64),
                ("n_out", c_size_t), ("out", _U64P), ("n_trace", c_size_t),
                ("trace", POINTER(TraceRec)), ("n_trace_init", c_size_t),
                ("trace_init", POINTER(TraceRec)), ("sizes", c_uint64 * 4),
                ("extra", _U64P * 3), ("n_extra", c_size_t * 3),
                ("n_oplog", c_size_t), ("oplog", POINTER(OpRec))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError(f"{LIB} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB)
        L.orc_sort_unique.restype = c_size_t
        L.orc_sort_unique.argtypes = [_U64P, c_size_t]
        L.orc_merge_union.restype = c_size_t
        L.orc_merge_union.argtypes = [_U64P, c_size_t, _U64P, c_size_t, _U64P]
        L.orc_set_difference.restype = c_size_t
        L.orc_set_difference.argtypes = [_U64P, c_size_t, _U64P, c_size_t, _U64P]
        L.orc_tc.restype = c_int
        L.orc_tc.argtypes = [_U64P, c_size_t, c_int, POINTER(Result)]
        L.orc_sg.restype = c_int
        L.orc_sg.argtypes = [_U64P, c_size_t, POINTER(Result)]
        L.orc_cspa.restype = c_int
        L.orc_cspa.argtypes = [_U64P, c_size_t, _U64P, c_size_t, POINTER(Result)]
        L.orc_result_free.restype = None
        L.orc_result_free.argtypes = [POINTER(Result)]
        for fn in (L.souffle_sg, L.souffle_tc):
            fn.restype = c_int
            fn.argtypes = [_U64P, c_uint64, c_int, c_int, _U64P, _U64P, POINTER(c_int), _U64P,
                           c_int]
        L.souffle_cspa.restype = c_int
        L.souffle_cspa.argtypes = [_U64P, c_uint64, _U64P, c_uint64, c_int, c_int, _U64P, _U64P,
                                   POINTER(c_int)]
        L.cpu_tc_omp.restype = c_int
        L.cpu_tc_omp.argtypes = [_U64P, c_uint64, c_int, c_int, _U64P, _U64P,
                                 POINTER(c_int), _U64P, c_int]
        _lib = L
    return _lib


def tc_omp(edges, threads, max_iteration=-1, trace_cap=1 << 20):
    """Multi-core CPU baseline of the TC loop (cpu_tc_omp.c): |path|,
    derived, iterations and per-body |delta| (no tuples)."""
    e = _arr(edges)
    tr = np.zeros(trace_cap, dtype=np.uint64)
    n_out, der = c_uint64(), c_uint64()
    it = c_int()
    if lib().cpu_tc_omp(_p(e), e.size, threads, max_iteration, ctypes.byref(n_out),
                        ctypes.byref(der), ctypes.byref(it), _p(tr), trace_cap) != 0:
        raise MemoryError("cpu_tc_omp")
    return dict(n_out=n_out.value, derived=der.value, iterations=it.value,
                deltas=[int(x) for x in tr[:min(it.value, trace_cap)]])


def _souffle(fn, edges, threads, max_iteration, trace_cap):
    e = _arr(edges)
    tr = np.zeros(trace_cap, dtype=np.uint64)
    n_out, cs = c_uint64(), c_uint64()
    it = c_int()
    if fn(_p(e), e.size, threads, max_iteration, ctypes.byref(n_out), ctypes.byref(cs),
          ctypes.byref(it), _p(tr), trace_cap) != 0:
        raise MemoryError(fn.__name__)
    return dict(n_out=n_out.value, checksum=cs.value, iterations=it.value, trace=tr)


def souffle_sg(edges, threads, max_iteration=-1, trace_cap=1 << 16):
    """Multi-core restatement of Souffle's SG plan (cpu_souffle.c): |sg|,
    checksum, recursive bodies, per-body |new| (entry 0: the x != y rule)."""
    r = _souffle(lib().souffle_sg, edges, threads, max_iteration, trace_cap)
    r["trace"] = [int(x) for x in r["trace"][:min(r["iterations"] + 1, trace_cap)]]
    return r


def souffle_tc(edges, threads, max_iteration=-1, trace_cap=1 << 16):
    """Multi-core restatement of Souffle's non-linear TC plan
    (cpu_souffle.c): |path|, checksum, recursive bodies, per-body |new|."""
    r = _souffle(lib().souffle_tc, edges, threads, max_iteration, trace_cap)
    r["trace"] = [int(x) for x in r["trace"][:min(r["iterations"], trace_cap)]]
    return r


def souffle_cspa(assign, deref, threads, max_iteration=-1):
    """Multi-core restatement of Souffle's CSPA plan (cpu_souffle.c): sizes
    and checksums of ValueFlow, MemoryAlias, ValueAlias; recursive bodies."""
    a, d = _arr(assign), _arr(deref)
    sz = np.zeros(3, dtype=np.uint64)
    cs = np.zeros(3, dtype=np.uint64)
    it = c_int()
    if lib().souffle_cspa(_p(a), a.size, _p(d), d.size, threads, max_iteration, _p(sz), _p(cs),
                          ctypes.byref(it)) != 0:
        raise MemoryError("souffle_cspa")
    return dict(sizes=[int(x) for x in sz], checksums=[int(x) for x in cs],
                iterations=it.value)


def _p(a):
    return a.ctypes.data_as(_U64P)


def _arr(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def sort_unique(keys):
    a = _arr(keys).copy()
    n = lib().orc_sort_unique(_p(a), a.size)
    return a[:n]


def merge_union(a, b):
    a, b = _arr(a), _arr(b)
    out = np.zeros(a.size + b.size + 1, dtype=np.uint64)
    n = lib().orc_merge_union(_p(a), a.size, _p(b), b.size, _p(out))
    return out[:n]


def set_difference(b, a):
    a, b = _arr(a), _arr(b)
    out = np.zeros(b.size + 1, dtype=np.uint64)
    n = lib().orc_set_difference(_p(b), b.size, _p(a), a.size, _p(out))
    return out[:n]


def _collect(res):
    out = np.ctypeslib.as_array(res.out, shape=(res.n_out,)).copy() if res.n_out else \
        np.zeros(0, dtype=np.uint64)
    tr = [(res.trace[i].iteration, res.trace[i].rel, res.trace[i].newt, res.trace[i].delta,
           res.trace[i].full) for i in range(res.n_trace)]
    ti = [(res.trace_init[i].iteration, res.trace_init[i].rel, res.trace_init[i].newt,
           res.trace_init[i].delta, res.trace_init[i].full) for i in range(res.n_trace_init)]
    extra = []
    for k in range(2):
        n = res.n_extra[k]
        extra.append(np.ctypeslib.as_array(res.extra[k], shape=(n,)).copy() if n else
                     np.zeros(0, dtype=np.uint64))
    ops = [(res.oplog[i].iteration, res.oplog[i].op, res.oplog[i].ran, res.oplog[i].raw,
            res.oplog[i].out_newt) for i in range(res.n_oplog)]
    r = dict(iterations=res.iterations, init_iterations=res.init_iterations,
             derived=res.derived, n_out=res.n_out, out=out, trace=tr, trace_init=ti,
             sizes=list(res.sizes), extra=extra, oplog=ops)
    lib().orc_result_free(ctypes.byref(res))
    return r


def tc(edges, max_iteration=-1):
    e = _arr(edges)
    res = Result()
    lib().orc_tc(_p(e), e.size, max_iteration, ctypes.byref(res))
    return _collect(res)


def path_length(edges):
    """All-pairs hop counts: restates test/path_length.cu:49-124 with numpy.

    path(a, c, l + 1) :- edge(a, b), path(b, c, l), path loaded from the edges
    with l = 1 (path_length.cu:83-87, 104-105); the set difference compares
    (from, to) only (lie.cu:131-138 over arity - dependent columns), so a
    pair keeps the l of the first iteration deriving it.  The first delta is
    the first |FULL| raw rows, as in TC (relation.cu:764-795 via lie.cu:56).
    Returns (sorted keys, hop counts aligned with them)."""
    e = _arr(edges)
    mask = np.uint64(0xffffffff)
    full = np.unique(e)
    # edges by target b: the inner edge_2__2_1 indexed on b
    by_b = full[np.argsort(full & mask, kind="stable")]
    tgt = by_b & mask
    src = by_b >> np.uint64(32)
    segments = [(full, 1)]
    delta = e[:full.size]
    length = 1
    while delta.size:
        b, c = delta >> np.uint64(32), delta & mask
        lo, hi = np.searchsorted(tgt, b, "left"), np.searchsorted(tgt, b, "right")
        cnt = hi - lo
        rows = np.repeat(np.arange(delta.size), cnt)
        pos = np.repeat(lo - np.cumsum(cnt) + cnt, cnt) + np.arange(rows.size)
        newt = np.unique((src[pos] << np.uint64(32)) | c[rows])
        fresh = np.setdiff1d(newt, full, assume_unique=True)
        length += 1
        if fresh.size:
            segments.append((fresh, length))
            full = np.union1d(full, fresh)
        delta = fresh
    keys = np.concatenate([s for s, _ in segments])
    lens = np.concatenate([np.full(s.size, l, dtype=np.uint32) for s, l in segments])
    order = np.argsort(keys, kind="stable")
    return keys[order], lens[order]


def sg(edges):
    e = _arr(edges)
    res = Result()
    lib().orc_sg(_p(e), e.size, ctypes.byref(res))
    return _collect(res)


def cspa(assign, deref):
    a, d = _arr(assign), _arr(deref)
    res = Result()
    lib().orc_cspa(_p(a), a.size, _p(d), d.size, ctypes.byref(res))
    return _collect(res)
