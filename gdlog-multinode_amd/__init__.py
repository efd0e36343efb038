"""ctypes bindings of the MI355X semi-naive Datalog engine.

The engine is the C ABI of ``include/gdl_hip.h`` (libgdl_hip.so) and the
program runners of ``include/gdl_programs.h`` (libgdlog.so, the gdlog C++ API
and the TC / SG / CSPA programs).  This module only marshals numpy arrays
through that ABI; it has no compute path of its own.  If the shared libraries
are missing, ``load()`` raises -- there is no fallback.

Seeded input generators (SplitMix64) come from libgdl_gen.so.
"""
import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int32,
                    c_uint32, c_uint64, c_void_p)

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
REPO = os.path.dirname(HERE)
INCLUDE_DIR = os.path.join(REPO, "include")

GDL_OK = 0
GDL_ERR_INVALID_ARG = 1
GDL_ERR_OVERFLOW = 5
STATUS_NAMES = {0: "ok", 1: "invalid argument", 2: "out of device memory", 3: "HIP error",
                4: "unsupported", 5: "size limit exceeded", 6: "device-side error",
                7: "busy"}

COL_INNER0, COL_INNER1, COL_OUTER0, COL_OUTER1 = 0, 1, 2, 3
PRED_NONE, PRED_EQ, PRED_NE, PRED_LT, PRED_LE, PRED_GT, PRED_GE, PRED_FALSE = range(8)


class GdlError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg}: {STATUS_NAMES.get(status, status)}")
        self.status = status


class Config(Structure):
    _fields_ = [("device", c_int), ("mem_cap_bytes", c_uint64), ("load_factor", c_float),
                ("collect_stats", c_int), ("verbose", c_int)]


class Index(Structure):
    _fields_ = [("slots", c_void_p), ("n_slots", c_uint64), ("n_distinct", c_uint64),
                ("prefix_cols", c_int), ("max_run", c_uint64), ("dense", c_int)]


class HashSet(Structure):
    _fields_ = [("slots", c_void_p), ("capacity", c_uint64), ("size", c_uint64),
                ("key_bits", c_uint32), ("pad_", c_uint32)]


FULL_MODE_AUTO, FULL_MODE_MERGE = 0, 1


def set_full_mode(mode):
    """FULL maintenance of later program runs (see gdl_prog_set_full_mode)."""
    _, prog = load()
    st = prog.gdl_prog_set_full_mode(mode)
    if st != GDL_OK:
        raise GdlError(st, "gdl_prog_set_full_mode")


def set_checkpoint(dir=None, every=0, resume=False, stop_after=-1):
    """Checkpoint / resume of later program runs on this thread (see
    gdl_prog_set_checkpoint); dir=None turns it off."""
    _, prog = load()
    d = None if dir is None else os.fsencode(dir)
    st = prog.gdl_prog_set_checkpoint(d, every, int(resume), stop_after)
    if st != GDL_OK:
        raise GdlError(st, "gdl_prog_set_checkpoint")


class ColMap(Structure):
    _fields_ = [("arity", c_int32), ("src", c_int32 * 2)]


class Pred(Structure):
    _fields_ = [("op", c_int32), ("lhs", c_int32), ("rhs", c_int32)]


class KernelStat(Structure):
    _fields_ = [("name", c_char_p), ("launches", c_uint64), ("total_ms", c_double),
                ("alg_bytes", c_double)]


class TraceRec(Structure):
    _fields_ = [("iteration", c_int32), ("relation", c_int32), ("newt", c_uint64),
                ("delta", c_uint64), ("full", c_uint64)]


class ProgResult(Structure):
    _fields_ = [("iterations", c_int32), ("init_iterations", c_int32), ("derived", c_uint64),
                ("n_out", c_uint64), ("sizes", c_uint64 * 4), ("load_ms", c_double),
                ("fixpoint_ms", c_double), ("n_trace", c_uint64), ("n_trace_init", c_uint64),
                ("checksum", c_uint64)]


class ProgBuffers(Structure):
    _fields_ = [("out", POINTER(c_uint64)), ("out_cap", c_uint64),
                ("extra0", POINTER(c_uint64)), ("extra0_cap", c_uint64),
                ("extra1", POINTER(c_uint64)), ("extra1_cap", c_uint64),
                ("trace", POINTER(TraceRec)), ("trace_cap", c_uint64),
                ("trace_init", POINTER(TraceRec)), ("trace_init_cap", c_uint64)]


class DenseRule(Structure):
    _fields_ = [("shape", c_int), ("r1", c_void_p), ("n1", c_uint64), ("r2", c_void_p),
                ("n2", c_uint64), ("flags", c_uint32), ("pad_", c_uint32)]


DENSE_TC, DENSE_SG = 1, 2
DENSE_LOG_LAST = 1


class HsSeg(Structure):
    _fields_ = [("slots", c_void_p), ("mask", c_uint64)]


class Tuning(Structure):
    _fields_ = [("load_factor", c_float), ("blocks_per_cu", c_uint32), ("radix_bits", c_uint32),
                ("join_block", c_uint32)]


class TcbConfig(Structure):
    _fields_ = [("max_slots", c_uint32), ("node_bits", c_uint32), ("mem_budget", c_uint64),
                ("keep_keys", c_int), ("solo_tuples", c_uint64), ("solo_threads", c_uint32)]


class TcbResult(Structure):
    _fields_ = [("qid", c_uint64), ("n_out", c_uint64), ("checksum", c_uint64),
                ("derived", c_uint64), ("iterations", c_int32), ("slot", c_int32),
                ("n_trace", c_uint64), ("admit_ms", c_double), ("done_ms", c_double),
                ("has_keys", c_int)]


GDL_ERR_BUSY = 7


_U64P = POINTER(c_uint64)
_engine = None
_programs = None
_gen = None


def lib_path(name):
    return os.path.join(LIB_DIR, name)


def _decl(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def load():
    """Loads libgdl_hip.so and libgdlog.so (raises OSError if not built)."""
    global _engine, _programs
    if _engine is not None:
        return _engine, _programs
    for n in ("libgdl_hip.so", "libgdlog.so"):
        if not os.path.exists(lib_path(n)):
            raise OSError(f"{lib_path(n)} is not built (run __graft_entry__.build())")
    eng = ctypes.CDLL(lib_path("libgdl_hip.so"), mode=ctypes.RTLD_GLOBAL)
    prog = ctypes.CDLL(lib_path("libgdlog.so"), mode=ctypes.RTLD_GLOBAL)
    vp = c_void_p
    _decl(eng, "gdl_config_default", None, [POINTER(Config)])
    _decl(eng, "gdl_ctx_create", c_int, [POINTER(Config), POINTER(vp)])
    _decl(eng, "gdl_ctx_destroy", c_int, [vp])
    _decl(eng, "gdl_status_string", c_char_p, [c_int])
    _decl(eng, "gdl_last_error", c_char_p, [vp])
    _decl(eng, "gdl_sync", c_int, [vp])
    _decl(eng, "gdl_stream", vp, [vp])
    _decl(eng, "gdl_device", c_int, [vp])
    _decl(eng, "gdl_set_load_factor", c_int, [vp, c_float])
    _decl(eng, "gdl_alloc", c_int, [vp, c_uint64, POINTER(vp)])
    _decl(eng, "gdl_free", c_int, [vp, vp])
    _decl(eng, "gdl_memcpy_h2d", c_int, [vp, vp, vp, c_uint64])
    _decl(eng, "gdl_memcpy_d2h", c_int, [vp, vp, vp, c_uint64])
    _decl(eng, "gdl_memcpy_d2d", c_int, [vp, vp, vp, c_uint64])
    _decl(eng, "gdl_mem_info", c_int, [vp] + [_U64P] * 5)
    _decl(eng, "gdl_release_cached", c_int, [vp])
    _decl(eng, "gdl_set_mem_cap", c_int, [vp, c_uint64])
    _decl(eng, "gdl_set_cache_limit", c_int, [vp, c_uint64])
    _decl(eng, "gdl_sort_u64", c_int, [vp, vp, c_uint64, c_int, c_int, vp])
    _decl(eng, "gdl_unique_u64", c_int, [vp, vp, c_uint64, vp, _U64P])
    _decl(eng, "gdl_key_bits", c_int, [vp, vp, c_uint64, POINTER(c_int), POINTER(c_int)])
    _decl(eng, "gdl_sort_unique_u64", c_int, [vp, vp, c_uint64, c_int, c_int, vp, _U64P])
    _decl(eng, "gdl_hashset_create_bitmap", c_int, [vp, c_uint32, vp])
    _decl(eng, "gdl_set_id_bits", c_int, [vp, c_int])
    _decl(eng, "gdl_hashset_check", c_int, [vp, vp])
    _decl(eng, "gdl_sort_unique_u64_noalloc", c_int,
          [vp, vp, c_uint64, c_int, c_int, vp, _U64P])
    _decl(eng, "gdl_merge_union_u64", c_int, [vp, vp, c_uint64, vp, c_uint64, vp, _U64P])
    _decl(eng, "gdl_set_difference_u64", c_int, [vp, vp, c_uint64, vp, c_uint64, vp, _U64P])
    _decl(eng, "gdl_merge_diff_u64", c_int,
          [vp, vp, c_uint64, vp, c_uint64, vp, _U64P, vp, _U64P])
    _decl(eng, "gdl_index_build", c_int, [vp, vp, c_uint64, c_int, POINTER(Index)])
    _decl(eng, "gdl_index_free", c_int, [vp, POINTER(Index)])
    _decl(eng, "gdl_index_lookup", c_int,
          [vp, POINTER(Index), vp, c_uint64, vp, c_uint64, vp, vp])
    _decl(eng, "gdl_join_count", c_int,
          [vp, vp, c_uint64, POINTER(Index), vp, c_uint64, vp, vp])
    _decl(eng, "gdl_join_scan", c_int, [vp, vp, c_uint64, vp, _U64P])
    _decl(eng, "gdl_join_materialize", c_int,
          [vp, vp, vp, c_uint64, vp, vp, c_uint64, ColMap, vp])
    _decl(eng, "gdl_join", c_int,
          [vp, vp, c_uint64, POINTER(Index), vp, c_uint64, ColMap, Pred, POINTER(vp), _U64P])
    _decl(eng, "gdl_project_u64", c_int, [vp, vp, c_uint64, ColMap, vp])
    _decl(eng, "gdl_filter_u64", c_int, [vp, vp, c_uint64, Pred, vp, _U64P])
    _decl(eng, "gdl_stats_enable", c_int, [vp, c_int])
    _decl(eng, "gdl_stats_reset", c_int, [vp])
    _decl(eng, "gdl_stats_get", c_int, [vp, POINTER(KernelStat), c_int, POINTER(c_int)])
    rp, bp = POINTER(ProgResult), POINTER(ProgBuffers)
    _decl(prog, "gdl_prog_tc", c_int, [vp, vp, c_uint64, c_int, c_int, rp, bp])
    _decl(prog, "gdl_prog_sg", c_int, [vp, vp, c_uint64, c_int, rp, bp])
    _decl(prog, "gdl_prog_tc_hooks", c_int, [vp, vp, c_uint64, c_int, c_int, rp, bp])
    _decl(prog, "gdl_prog_cspa", c_int, [vp, vp, c_uint64, vp, c_uint64, c_int, rp, bp])
    _decl(prog, "gdl_prog_last_error", c_char_p, [])
    _decl(prog, "gdl_prog_path_length", c_int,
          [vp, vp, c_uint64, c_int, rp, vp, vp, c_uint64])
    _decl(prog, "gdl_prog_path_length_hooks", c_int,
          [vp, vp, c_uint64, c_int, rp, vp, vp, c_uint64])
    _decl(prog, "gdl_prog_set_full_mode", c_int, [c_int])
    _decl(prog, "gdl_prog_set_checkpoint", c_int, [ctypes.c_char_p, c_int, c_int, c_int])
    _decl(prog, "gdl_datalog_compile", c_int, [c_char_p, POINTER(vp)])
    _decl(prog, "gdl_datalog_compile_slog", c_int, [c_char_p, POINTER(vp)])
    _decl(prog, "gdl_datalog_source", c_char_p, [vp])
    _decl(prog, "gdl_datalog_destroy", None, [vp])
    _decl(prog, "gdl_datalog_plan", c_char_p, [vp])
    _decl(prog, "gdl_datalog_run", c_int,
          [vp, vp, c_uint32, POINTER(c_char_p), POINTER(_U64P), _U64P])
    _decl(prog, "gdl_datalog_result", c_int, [vp, c_char_p, _U64P, c_uint64, _U64P])
    _decl(prog, "gdl_datalog_stats", c_int,
          [vp, _U64P, POINTER(c_int32), c_uint32, POINTER(c_uint32)])
    _decl(prog, "gdl_datalog_kept", c_char_p, [vp, c_uint32])
    _decl(prog, "gdl_datalog_arity", c_int, [vp, c_char_p])
    _decl(prog, "gdl_datalog_last_error", c_char_p, [])
    _decl(eng, "gdl_hashset_create", c_int, [vp, c_uint64, POINTER(HashSet)])
    _decl(eng, "gdl_hashset_free", c_int, [vp, POINTER(HashSet)])
    _decl(eng, "gdl_hashset_insert_new", c_int,
          [vp, POINTER(HashSet), vp, c_uint64, vp, _U64P])
    _decl(eng, "gdl_join_insert", c_int,
          [vp, vp, c_uint64, POINTER(Index), vp, c_uint64, ColMap, Pred, POINTER(HashSet),
           POINTER(vp), _U64P, _U64P, _U64P])
    _decl(eng, "gdl_get_tuning", c_int, [vp, POINTER(Tuning)])
    _decl(eng, "gdl_set_tuning", c_int, [vp, POINTER(Tuning)])
    _decl(eng, "gdl_hashset_reserve", c_int, [vp, POINTER(HashSet), c_uint64])
    _decl(eng, "gdl_hashset_extract", c_int, [vp, POINTER(HashSet), vp, c_uint64, _U64P])
    _decl(eng, "gdl_hashset_extract_sorted", c_int, [vp, POINTER(HashSet), vp, c_uint64, _U64P])
    _decl(eng, "gdl_set_dense_mode", c_int, [vp, c_int])
    _decl(eng, "gdl_fixpoint_dense", c_int,
          [vp, POINTER(DenseRule), POINTER(HashSet), vp, c_uint64, POINTER(vp), _U64P, _U64P,
           c_int, _U64P, _U64P, POINTER(c_int)])
    _decl(eng, "gdl_join_insert_seg", c_int,
          [vp, vp, c_uint64, POINTER(Index), vp, c_uint64, ColMap, Pred, vp, c_uint32, c_int,
           vp, c_uint64, _U64P, _U64P])
    _decl(eng, "gdl_key_histogram_u64", c_int, [vp, vp, c_uint64, c_int, c_uint32, vp])
    _decl(eng, "gdl_mask_columns_u64", c_int, [vp, vp, c_uint64, c_uint32, vp])
    _decl(eng, "gdl_segment_labels_u32", c_int, [vp, vp, c_uint32, c_uint32, vp, c_uint64])
    _decl(eng, "gdl_segment_labels_sorted", c_int,
          [vp, vp, c_uint64, vp, c_uint32, c_uint32, c_int, vp, vp])
    _decl(eng, "gdl_graph_begin", c_int, [vp])
    _decl(eng, "gdl_graph_end", c_int, [vp, POINTER(vp)])
    _decl(eng, "gdl_graph_launch", c_int, [vp, vp])
    _decl(eng, "gdl_graph_destroy", c_int, [vp])
    _decl(eng, "gdl_tcb_config_default", None, [POINTER(TcbConfig)])
    _decl(eng, "gdl_tcb_create", c_int, [vp, POINTER(TcbConfig), POINTER(vp)])
    _decl(eng, "gdl_tcb_destroy", c_int, [vp])
    _decl(eng, "gdl_tc_estimate", c_uint64, [vp, c_uint64, c_uint32, c_uint32, c_uint64])
    _decl(eng, "gdl_tcb_admit", c_int, [vp, c_uint64, vp, c_uint64, c_uint32, c_uint64])
    _decl(eng, "gdl_tcb_step", c_int, [vp, POINTER(c_uint32)])
    _decl(eng, "gdl_tcb_poll", c_int, [vp, POINTER(TcbResult)])
    _decl(eng, "gdl_tcb_fetch", c_int, [vp, c_uint64, vp, vp])
    _decl(eng, "gdl_tcb_info", c_int, [vp] + [_U64P] * 5)
    _engine, _programs = eng, prog
    return eng, prog


def gen_lib():
    """libgdl_gen.so (pure C, no GPU needed)."""
    global _gen
    if _gen is None:
        p = lib_path("libgdl_gen.so")
        if not os.path.exists(p):
            raise OSError(f"{p} is not built (run __graft_entry__.build())")
        g = ctypes.CDLL(p)
        _decl(g, "gdl_gen_er", c_uint64, [c_uint64, c_uint32, c_uint64, _U64P])
        _decl(g, "gdl_gen_tree", c_uint64, [c_uint64, c_uint32, c_uint32, _U64P])
        _decl(g, "gdl_gen_rmat", c_uint64,
              [c_uint64, c_int, c_uint64, c_double, c_double, c_double, _U64P])
        _decl(g, "gdl_gen_batched_shape", None,
              [c_uint64, c_uint64, c_uint64, POINTER(c_int), _U64P, POINTER(c_uint32)])
        _decl(g, "gdl_gen_batched_instance", c_uint64,
              [c_uint64, c_uint64, c_uint64, c_uint32, _U64P])
        _decl(g, "gdl_gen_cspa", c_uint64,
              [c_uint64, c_uint32, c_double, c_uint64, c_uint64, _U64P, _U64P, _U64P])
        _gen = g
    return _gen


def _p64(a):
    return a.ctypes.data_as(_U64P)


# --------------------------------------------------------------- generators

def gen_er(seed, n_nodes, m):
    out = np.zeros(m, dtype=np.uint64)
    k = gen_lib().gdl_gen_er(seed, n_nodes, m, _p64(out))
    return out[:k]


def gen_tree(seed, n_nodes, window=1000):
    out = np.zeros(max(n_nodes - 1, 1), dtype=np.uint64)
    k = gen_lib().gdl_gen_tree(seed, n_nodes, window, _p64(out))
    return out[:k]


def gen_rmat(seed, scale, m, a=0.57, b=0.19, c=0.19):
    out = np.zeros(m, dtype=np.uint64)
    k = gen_lib().gdl_gen_rmat(seed, scale, m, a, b, c, _p64(out))
    return out[:k]


def gen_batched_shape(seed, min_edges=1000, max_edges=50000):
    t, m, n = c_int(), c_uint64(), c_uint32()
    gen_lib().gdl_gen_batched_shape(seed, min_edges, max_edges, ctypes.byref(t),
                                    ctypes.byref(m), ctypes.byref(n))
    return bool(t.value), int(m.value), int(n.value)


def gen_batched_instance(seed, min_edges=1000, max_edges=50000, tree_window=1000):
    _, m, _ = gen_batched_shape(seed, min_edges, max_edges)
    out = np.zeros(m + 1, dtype=np.uint64)
    k = gen_lib().gdl_gen_batched_instance(seed, min_edges, max_edges, tree_window, _p64(out))
    return out[:k]


def gen_cspa(seed, n_ids, zipf_s, n_assign, n_deref):
    a = np.zeros(n_assign, dtype=np.uint64)
    d = np.zeros(n_deref, dtype=np.uint64)
    nd = c_uint64()
    na = gen_lib().gdl_gen_cspa(seed, n_ids, zipf_s, n_assign, n_deref, _p64(a), _p64(d),
                                ctypes.byref(nd))
    return a[:na], d[:nd.value]


def pack(src, dst):
    return (np.asarray(src, dtype=np.uint64) << np.uint64(32)) | np.asarray(dst, dtype=np.uint64)


def unpack(keys):
    k = np.asarray(keys, dtype=np.uint64)
    return (k >> np.uint64(32)).astype(np.int64), (k & np.uint64(0xffffffff)).astype(np.int64)


# ------------------------------------------------------------------ engine

class DeviceArray:
    """A device buffer of u64 (or u32) owned by a Context's allocator."""

    def __init__(self, ctx, ptr, n, dtype=np.uint64):
        self.ctx, self.ptr, self.n, self.dtype = ctx, ptr, int(n), np.dtype(dtype)

    def numpy(self, n=None):
        n = self.n if n is None else int(n)
        out = np.zeros(n, dtype=self.dtype)
        if n:
            self.ctx._check(self.ctx.eng.gdl_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr,
                                                        n * self.dtype.itemsize), "d2h")
        return out

    def free(self):
        if self.ptr:
            self.ctx.eng.gdl_free(self.ctx.h, self.ptr)
            self.ptr = None


class Context:
    def __init__(self, device=0, mem_cap_bytes=0, load_factor=0.35, collect_stats=False,
                 verbose=False):
        self.eng, self.prog = load()
        cfg = Config()
        self.eng.gdl_config_default(ctypes.byref(cfg))
        cfg.device, cfg.mem_cap_bytes, cfg.load_factor = device, mem_cap_bytes, load_factor
        cfg.collect_stats, cfg.verbose = int(collect_stats), int(verbose)
        h = c_void_p()
        st = self.eng.gdl_ctx_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != GDL_OK:
            raise GdlError(st, "gdl_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            self.eng.gdl_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, what):
        if st != GDL_OK:
            detail = self.eng.gdl_last_error(self.h) or b""
            raise GdlError(st, f"{what} ({detail.decode(errors='replace')})")

    # memory
    def alloc(self, n, dtype=np.uint64):
        p = c_void_p()
        self._check(self.eng.gdl_alloc(self.h, max(int(n), 1) * np.dtype(dtype).itemsize,
                                       ctypes.byref(p)), "gdl_alloc")
        return DeviceArray(self, p.value, n, dtype)

    def to_device(self, arr, dtype=np.uint64):
        a = np.ascontiguousarray(arr, dtype=dtype)
        d = self.alloc(a.size, dtype)
        if a.size:
            self._check(self.eng.gdl_memcpy_h2d(self.h, d.ptr, a.ctypes.data, a.nbytes), "h2d")
        return d

    def sync(self):
        self._check(self.eng.gdl_sync(self.h), "gdl_sync")

    def mem_info(self):
        v = [c_uint64() for _ in range(5)]
        self._check(self.eng.gdl_mem_info(self.h, *[ctypes.byref(x) for x in v]), "mem_info")
        return dict(zip(("free", "total", "in_use", "peak", "cached"), [x.value for x in v]))

    def set_mem_cap(self, cap):
        self._check(self.eng.gdl_set_mem_cap(self.h, cap), "set_mem_cap")

    def set_cache_limit(self, nbytes):
        self._check(self.eng.gdl_set_cache_limit(self.h, nbytes), "set_cache_limit")

    def tuning(self):
        t = Tuning()
        self._check(self.eng.gdl_get_tuning(self.h, ctypes.byref(t)), "get_tuning")
        return dict(load_factor=round(float(t.load_factor), 4), blocks_per_cu=t.blocks_per_cu,
                    radix_bits=t.radix_bits, join_block=t.join_block)

    def set_tuning(self, load_factor=None, blocks_per_cu=None, radix_bits=None,
                   join_block=None):
        cur = self.tuning()
        t = Tuning(cur["load_factor"] if load_factor is None else load_factor,
                   cur["blocks_per_cu"] if blocks_per_cu is None else blocks_per_cu,
                   cur["radix_bits"] if radix_bits is None else radix_bits,
                   cur["join_block"] if join_block is None else join_block)
        self._check(self.eng.gdl_set_tuning(self.h, ctypes.byref(t)), "set_tuning")

    def release_cached(self):
        self._check(self.eng.gdl_release_cached(self.h), "release_cached")

    # primitives (numpy in, numpy out)
    def sort(self, keys, hi_bits=32, lo_bits=32):
        d = self.to_device(keys)
        o = self.alloc(d.n)
        self._check(self.eng.gdl_sort_u64(self.h, d.ptr, d.n, hi_bits, lo_bits, o.ptr), "sort")
        r = o.numpy()
        d.free(); o.free()
        return r

    def sort_unique(self, keys, hi_bits=32, lo_bits=32):
        d = self.to_device(keys)
        o = self.alloc(d.n)
        m = c_uint64()
        self._check(self.eng.gdl_sort_unique_u64(self.h, d.ptr, d.n, hi_bits, lo_bits, o.ptr,
                                                 ctypes.byref(m)), "sort_unique")
        r = o.numpy(m.value)
        d.free(); o.free()
        return r

    def unique(self, sorted_keys):
        d = self.to_device(sorted_keys)
        o = self.alloc(d.n)
        m = c_uint64()
        self._check(self.eng.gdl_unique_u64(self.h, d.ptr, d.n, o.ptr, ctypes.byref(m)), "unique")
        r = o.numpy(m.value)
        d.free(); o.free()
        return r

    def key_bits(self, keys):
        d = self.to_device(keys)
        hi, lo = c_int(), c_int()
        self._check(self.eng.gdl_key_bits(self.h, d.ptr, d.n, ctypes.byref(hi), ctypes.byref(lo)),
                    "key_bits")
        d.free()
        return hi.value, lo.value

    def merge_union(self, a, b):
        da, db = self.to_device(a), self.to_device(b)
        o = self.alloc(da.n + db.n)
        m = c_uint64()
        self._check(self.eng.gdl_merge_union_u64(self.h, da.ptr, da.n, db.ptr, db.n, o.ptr,
                                                 ctypes.byref(m)), "merge_union")
        r = o.numpy(m.value)
        for x in (da, db, o):
            x.free()
        return r

    def set_difference(self, b, a):
        da, db = self.to_device(a), self.to_device(b)
        o = self.alloc(db.n)
        m = c_uint64()
        self._check(self.eng.gdl_set_difference_u64(self.h, db.ptr, db.n, da.ptr, da.n, o.ptr,
                                                    ctypes.byref(m)), "set_difference")
        r = o.numpy(m.value)
        for x in (da, db, o):
            x.free()
        return r

    def merge_diff(self, full, newt):
        df, dn = self.to_device(full), self.to_device(newt)
        u, d = self.alloc(df.n + dn.n), self.alloc(dn.n)
        nu, nd = c_uint64(), c_uint64()
        self._check(self.eng.gdl_merge_diff_u64(self.h, df.ptr, df.n, dn.ptr, dn.n, u.ptr,
                                                ctypes.byref(nu), d.ptr, ctypes.byref(nd)),
                    "merge_diff")
        ru, rd = u.numpy(nu.value), d.numpy(nd.value)
        for x in (df, dn, u, d):
            x.free()
        return ru, rd

    def index_build(self, dkeys, prefix_cols=1):
        idx = Index()
        self._check(self.eng.gdl_index_build(self.h, dkeys.ptr, dkeys.n, prefix_cols,
                                             ctypes.byref(idx)), "index_build")
        return idx

    def index_free(self, idx):
        self._check(self.eng.gdl_index_free(self.h, ctypes.byref(idx)), "index_free")

    def index_lookup(self, idx, dkeys, prefixes):
        dq = self.to_device(prefixes)
        b, c = self.alloc(dq.n), self.alloc(dq.n, np.uint32)
        self._check(self.eng.gdl_index_lookup(self.h, ctypes.byref(idx), dkeys.ptr, dkeys.n,
                                              dq.ptr, dq.n, b.ptr, c.ptr), "index_lookup")
        rb, rc = b.numpy(), c.numpy()
        for x in (dq, b, c):
            x.free()
        return rb, rc

    def join(self, inner_sorted, outer, colmap, prefix_cols=1, pred=(PRED_NONE, 0, 1)):
        """Hash join of sorted-unique inner keys with outer keys (any order)."""
        di, do = self.to_device(inner_sorted), self.to_device(outer)
        idx = self.index_build(di, prefix_cols)
        cm = ColMap(len(colmap), (c_int32 * 2)(*(list(colmap) + [0])[:2]))
        pr = Pred(*pred)
        out, m = c_void_p(), c_uint64()
        self._check(self.eng.gdl_join(self.h, di.ptr, di.n, ctypes.byref(idx), do.ptr, do.n, cm,
                                      pr, ctypes.byref(out), ctypes.byref(m)), "join")
        res = DeviceArray(self, out.value, m.value).numpy() if m.value else \
            np.zeros(0, dtype=np.uint64)
        if out.value:
            self.eng.gdl_free(self.h, out.value)
        self.index_free(idx)
        di.free(); do.free()
        return res

    def hashset_create(self, expected=0):
        hs = HashSet()
        self._check(self.eng.gdl_hashset_create(self.h, expected, ctypes.byref(hs)),
                    "hashset_create")
        return hs

    def hashset_create_bitmap(self, key_bits):
        hs = HashSet()
        self._check(self.eng.gdl_hashset_create_bitmap(self.h, key_bits, ctypes.byref(hs)),
                    "hashset_create_bitmap")
        return hs

    def hashset_check(self, hs):
        """Status of gdl_hashset_check (0, or the overflow status of a bitmap
        set that was offered a key outside its domain)."""
        return self.eng.gdl_hashset_check(self.h, ctypes.byref(hs))

    def hashset_free(self, hs):
        self._check(self.eng.gdl_hashset_free(self.h, ctypes.byref(hs)), "hashset_free")

    def hashset_insert_new(self, hs, keys):
        """Inserts keys; returns the keys that were not yet members (any order)."""
        d = self.to_device(keys)
        o = self.alloc(d.n)
        m = c_uint64()
        self._check(self.eng.gdl_hashset_insert_new(self.h, ctypes.byref(hs), d.ptr, d.n, o.ptr,
                                                    ctypes.byref(m)), "hashset_insert_new")
        r = o.numpy(m.value)
        d.free(); o.free()
        return r

    def join_insert(self, hs, inner_sorted, outer, colmap, prefix_cols=1, pred=(PRED_NONE, 0, 1)):
        """Fused join + insert-if-absent; returns (new keys in any order, produced count)."""
        di, do = self.to_device(inner_sorted), self.to_device(outer)
        idx = self.index_build(di, prefix_cols)
        cm = ColMap(len(colmap), (c_int32 * 2)(*(list(colmap) + [0])[:2]))
        buf, cap, nn, raw = c_void_p(), c_uint64(0), c_uint64(0), c_uint64(0)
        self._check(self.eng.gdl_join_insert(self.h, di.ptr, di.n, ctypes.byref(idx), do.ptr,
                                             do.n, cm, Pred(*pred), ctypes.byref(hs),
                                             ctypes.byref(buf), ctypes.byref(cap),
                                             ctypes.byref(nn), ctypes.byref(raw)), "join_insert")
        new = DeviceArray(self, buf.value, nn.value).numpy() if nn.value else \
            np.zeros(0, dtype=np.uint64)
        if buf.value:
            self.eng.gdl_free(self.h, buf.value)
        self.index_free(idx)
        di.free(); do.free()
        return new, raw.value

    def set_dense_mode(self, mode):
        """0: no bit-matrix bodies; 1: when the cost model prefers them (default);
        2: whenever they apply; 3 / 4: as 2, with SG's second join always / never
        expanded by columns (two bit-matrix transposes around a row OR)."""
        self._check(self.eng.gdl_set_dense_mode(self.h, int(mode)), "set_dense_mode")

    def hashset_extract_sorted(self, hs):
        """Members of a set in key order."""
        o = self.alloc(max(hs.size, 1))
        m = c_uint64()
        self._check(self.eng.gdl_hashset_extract_sorted(
            self.h, ctypes.byref(hs), o.ptr, max(hs.size, 1), ctypes.byref(m)),
            "hashset_extract_sorted")
        r = o.numpy(m.value)
        o.free()
        return r

    def fixpoint_dense(self, shape, key_bits, full, delta, r1, r2=None, max_bodies=1 << 20,
                       flags=0):
        """gdl_fixpoint_dense from host arrays: FULL (a set; also the log's start), the
        delta, the static relations.  Returns dict(new, raw, log, members) with the
        per-body counts, the log (FULL then each body's new tuples) and the bitmap set's
        members in key order afterwards."""
        hs = self.hashset_create_bitmap(key_bits)
        df = self.to_device(full)
        log = c_void_p(None)
        cap = c_uint64(0)
        end = c_uint64(int(df.n))
        bufs = []
        try:
            if df.n:
                scratch = self.alloc(df.n)
                added = c_uint64()
                self._check(self.eng.gdl_hashset_insert_new(
                    self.h, ctypes.byref(hs), df.ptr, df.n, scratch.ptr, ctypes.byref(added)),
                    "hashset_insert_new")
                scratch.free()
                p = c_void_p()
                self._check(self.eng.gdl_alloc(self.h, df.n * 8, ctypes.byref(p)), "alloc")
                self._check(self.eng.gdl_memcpy_d2d(self.h, p, df.ptr, df.n * 8), "d2d")
                log, cap = p, c_uint64(int(df.n))
            dd, d1 = self.to_device(delta), self.to_device(r1)
            bufs += [dd, d1]
            # the same array for r1 and r2 passes one device buffer (one relation)
            d2 = d1 if r2 is r1 else (self.to_device(r2) if r2 is not None else None)
            if d2 is not None and d2 is not d1:
                bufs.append(d2)
            rule = DenseRule(shape, d1.ptr, d1.n, d2.ptr if d2 is not None else None,
                             d2.n if d2 is not None else 0, flags, 0)
            hn = np.zeros(max_bodies, dtype=np.uint64)
            hr = np.zeros(max_bodies, dtype=np.uint64)
            nb = c_int()
            self._check(self.eng.gdl_fixpoint_dense(
                self.h, ctypes.byref(rule), ctypes.byref(hs), dd.ptr, dd.n, ctypes.byref(log),
                ctypes.byref(cap), ctypes.byref(end), max_bodies, _p64(hn), _p64(hr),
                ctypes.byref(nb)), "fixpoint_dense")
            out = np.zeros(end.value, dtype=np.uint64)
            if end.value:
                self._check(self.eng.gdl_memcpy_d2h(self.h, _p64(out), log, end.value * 8),
                            "d2h")
            members = self.hashset_extract_sorted(hs)
            return dict(new=[int(x) for x in hn[:nb.value]], raw=[int(x) for x in hr[:nb.value]],
                        log=out, members=members, size=int(hs.size))
        finally:
            if log.value:
                self.eng.gdl_free(self.h, log)
            for b in bufs:
                b.free()
            df.free()
            self.hashset_free(hs)

    def hashset_extract(self, hs):
        """Members of a set (any order)."""
        o = self.alloc(max(hs.size, 1))
        m = c_uint64()
        self._check(self.eng.gdl_hashset_extract(self.h, ctypes.byref(hs), o.ptr, max(hs.size, 1),
                                                 ctypes.byref(m)), "hashset_extract")
        r = o.numpy(m.value)
        o.free()
        return r

    def join_insert_seg(self, sets, seg_shift, inner_sorted, outer, colmap, prefix_cols=1,
                        pred=(PRED_NONE, 0, 1), out_cap=None):
        """Segmented fused join into per-segment sets (HashSet list, index = key >> seg_shift;
        None = no set).  The sets must already hold room; returns (new keys, produced)."""
        di, do = self.to_device(inner_sorted), self.to_device(outer)
        idx = self.index_build(di, prefix_cols)
        cm = ColMap(len(colmap), (c_int32 * 2)(*(list(colmap) + [0])[:2]))
        tab = (HsSeg * len(sets))()
        for i, hs in enumerate(sets):
            tab[i] = HsSeg(hs.slots, hs.capacity - 1) if hs is not None else HsSeg(None, 0)
        dt = self.alloc(2 * len(sets))
        self._check(self.eng.gdl_memcpy_h2d(self.h, dt.ptr, ctypes.addressof(tab),
                                            ctypes.sizeof(tab)), "h2d")
        cap = out_cap if out_cap is not None else max(1, do.n * max(1, idx.max_run))
        o = self.alloc(cap)
        nn, raw = c_uint64(), c_uint64()
        try:
            self._check(self.eng.gdl_join_insert_seg(
                self.h, di.ptr, di.n, ctypes.byref(idx), do.ptr, do.n, cm, Pred(*pred), dt.ptr,
                len(sets), seg_shift, o.ptr, cap, ctypes.byref(nn), ctypes.byref(raw)),
                "join_insert_seg")
            return o.numpy(nn.value), raw.value
        finally:
            self.index_free(idx)
            for a in (di, do, dt, o):
                a.free()

    def key_histogram(self, keys, shift, n_bins):
        d = self.to_device(keys)
        c = self.to_device(np.zeros(n_bins, dtype=np.uint64))
        self._check(self.eng.gdl_key_histogram_u64(self.h, d.ptr, d.n, shift, n_bins, c.ptr),
                    "key_histogram")
        r = c.numpy()
        d.free(); c.free()
        return r

    # graph capture of the context's stream (include/gdl_hip.h): between
    # graph_begin() and graph_end() only kernel-only entry points on
    # preallocated device buffers may be called
    def graph_begin(self):
        self._check(self.eng.gdl_graph_begin(self.h), "gdl_graph_begin")

    def graph_end(self):
        g = c_void_p()
        self._check(self.eng.gdl_graph_end(self.h, ctypes.byref(g)), "gdl_graph_end")
        return g

    def graph_launch(self, g):
        self._check(self.eng.gdl_graph_launch(self.h, g), "gdl_graph_launch")

    def graph_destroy(self, g):
        self.eng.gdl_graph_destroy(g)

    def mask_columns(self, keys, col_mask):
        d = self.to_device(keys)
        self._check(self.eng.gdl_mask_columns_u64(self.h, d.ptr, d.n, col_mask, d.ptr),
                    "mask_columns")
        r = d.numpy()
        d.free()
        return r

    def project(self, keys, colmap):
        d = self.to_device(keys)
        o = self.alloc(d.n)
        cm = ColMap(len(colmap), (c_int32 * 2)(*(list(colmap) + [0])[:2]))
        self._check(self.eng.gdl_project_u64(self.h, d.ptr, d.n, cm, o.ptr), "project")
        r = o.numpy()
        d.free(); o.free()
        return r

    def filter(self, keys, op, lhs=0, rhs=1):
        d = self.to_device(keys)
        o = self.alloc(d.n)
        m = c_uint64()
        self._check(self.eng.gdl_filter_u64(self.h, d.ptr, d.n, Pred(op, lhs, rhs), o.ptr,
                                            ctypes.byref(m)), "filter")
        r = o.numpy(m.value)
        d.free(); o.free()
        return r

    # statistics
    def stats_enable(self, on=True):
        self._check(self.eng.gdl_stats_enable(self.h, int(on)), "stats_enable")

    def stats_reset(self):
        self._check(self.eng.gdl_stats_reset(self.h), "stats_reset")

    def stats(self):
        arr = (KernelStat * 64)()
        n = c_int()
        self._check(self.eng.gdl_stats_get(self.h, arr, 64, ctypes.byref(n)), "stats_get")
        return {arr[i].name.decode(): dict(launches=arr[i].launches, total_ms=arr[i].total_ms,
                                           alg_bytes=arr[i].alg_bytes)
                for i in range(min(n.value, 64))}

    # programs
    def _run(self, fn, args, want_keys, want_trace, out_cap, extra_caps=(0, 0)):
        res = ProgResult()
        bufs = ProgBuffers()
        keep = []
        if want_keys:
            o = np.zeros(max(out_cap, 1), dtype=np.uint64)
            keep.append(o)
            bufs.out, bufs.out_cap = _p64(o), out_cap
            ex = []
            for i, cap in enumerate(extra_caps):
                e = np.zeros(max(cap, 1), dtype=np.uint64)
                ex.append(e)
                if cap:
                    if i == 0:
                        bufs.extra0, bufs.extra0_cap = _p64(e), cap
                    else:
                        bufs.extra1, bufs.extra1_cap = _p64(e), cap
            keep.append(ex)
        if want_trace:
            tr = (TraceRec * 100000)()
            ti = (TraceRec * 100000)()
            keep += [tr, ti]
            bufs.trace, bufs.trace_cap = tr, 100000
            bufs.trace_init, bufs.trace_init_cap = ti, 100000
        st = fn(self.h, *args, ctypes.byref(res), ctypes.byref(bufs))
        if st != GDL_OK:
            msg = (self.prog.gdl_prog_last_error() or b"").decode(errors="replace")
            raise GdlError(st, msg)
        r = dict(iterations=res.iterations, init_iterations=res.init_iterations,
                 derived=res.derived, n_out=res.n_out, sizes=list(res.sizes),
                 load_ms=res.load_ms, fixpoint_ms=res.fixpoint_ms, checksum=res.checksum)
        if want_keys:
            r["out"] = keep[0][:res.n_out].copy()
            r["extra"] = [keep[1][0][:res.sizes[0]].copy() if extra_caps[0] else None,
                          keep[1][1][:res.sizes[1]].copy() if extra_caps[1] else None]
        if want_trace:
            tr, ti = keep[-2], keep[-1]
            r["trace"] = [(tr[i].iteration, tr[i].relation, tr[i].newt, tr[i].delta, tr[i].full)
                          for i in range(res.n_trace)]
            r["trace_init"] = [(ti[i].iteration, ti[i].relation, ti[i].newt, ti[i].delta,
                                ti[i].full) for i in range(res.n_trace_init)]
        return r

    def _edges(self, edges, on_device):
        if on_device:
            if isinstance(edges, DeviceArray):
                return edges.ptr, edges.n, None
            d = self.to_device(edges)
            return d.ptr, d.n, d
        e = np.ascontiguousarray(edges, dtype=np.uint64)
        return e.ctypes.data, e.size, e

    def run_tc(self, edges, on_device=False, max_iteration=-1, want_keys=True, want_trace=True,
               out_cap=None):
        p, n, keep = self._edges(edges, on_device)
        cap = out_cap if out_cap is not None else (60_000_000 if want_keys else 0)
        try:
            return self._run(self.prog.gdl_prog_tc, (p, n, int(on_device), max_iteration),
                             want_keys, want_trace, cap)
        finally:
            if isinstance(keep, DeviceArray):
                keep.free()

    def run_tc_hooks(self, edges, variant, on_device=False, want_keys=True, want_trace=True,
                     out_cap=None):
        """TC with device-run hooks (gdl_prog_tc_hooks): variant 0 predicate
        t[1] != 42, 1 arithmetic generator, 2 arithmetic copy hook."""
        p, n, keep = self._edges(edges, on_device)
        cap = out_cap if out_cap is not None else (20_000_000 if want_keys else 0)
        try:
            return self._run(self.prog.gdl_prog_tc_hooks, (p, n, int(on_device), int(variant)),
                             want_keys, want_trace, cap)
        finally:
            if isinstance(keep, DeviceArray):
                keep.free()

    def run_path_length(self, edges, on_device=False, out_cap=None, hooks=False):
        """All-pairs hop counts (the reference's path_length program): returns
        the pairs (sorted keys), their hop counts, and the run's statistics.
        hooks=True runs the reference's own encoding (an arity-3 relation with
        a dependent column and an arithmetic generator hook run on the device)
        instead of the TC-based program."""
        p, n, keep = self._edges(edges, on_device)
        cap = out_cap if out_cap is not None else 60_000_000
        keys = np.empty(max(cap, 1), dtype=np.uint64)
        lens = np.empty(max(cap, 1), dtype=np.uint32)
        res = ProgResult()
        try:
            fn = self.prog.gdl_prog_path_length_hooks if hooks else self.prog.gdl_prog_path_length
            st = fn(self.h, p, n, int(on_device), ctypes.byref(res), keys.ctypes.data,
                    lens.ctypes.data, cap)
        finally:
            if isinstance(keep, DeviceArray):
                keep.free()
        if st != GDL_OK:
            msg = (self.prog.gdl_prog_last_error() or b"").decode(errors="replace")
            raise GdlError(st, msg)
        return dict(out=keys[:res.n_out].copy(), lens=lens[:res.n_out].copy(),
                    n_out=res.n_out, longest=res.sizes[0], iterations=res.iterations,
                    derived=res.derived, checksum=res.checksum, fixpoint_ms=res.fixpoint_ms)

    def run_sg(self, edges, on_device=False, want_keys=True, want_trace=True, out_cap=None):
        p, n, keep = self._edges(edges, on_device)
        cap = out_cap if out_cap is not None else (80_000_000 if want_keys else 0)
        try:
            return self._run(self.prog.gdl_prog_sg, (p, n, int(on_device)), want_keys,
                             want_trace, cap)
        finally:
            if isinstance(keep, DeviceArray):
                keep.free()

    def run_cspa(self, assign, deref, on_device=False, want_keys=True, want_trace=True,
                 out_cap=None, extra_cap=None):
        pa, na, ka = self._edges(assign, on_device)
        pd, nd, kd = self._edges(deref, on_device)
        cap = out_cap if out_cap is not None else (20_000_000 if want_keys else 0)
        ecap = extra_cap if extra_cap is not None else (5_000_000 if want_keys else 0)
        try:
            return self._run(self.prog.gdl_prog_cspa, (pa, na, pd, nd, int(on_device)),
                             want_keys, want_trace, cap, (ecap, ecap))
        finally:
            for k in (ka, kd):
                if isinstance(k, DeviceArray):
                    k.free()


# ------------------------------------------------------------ batched TC

class Datalog:
    """A Datalog program over unary and binary relations, compiled by the
    engine's front end (include/gdl_datalog.h) into LIEs of RA operators.
    Compiling needs no device; run() executes on a Context.  Relations hold
    packed keys (c0 << 32) | c1 (c0 << 32 for a unary relation).

        prog = Datalog(".decl edge(a: number, b: number) .input edge ..."
                       "path(x, y) :- edge(x, y). ...")
        out = prog.run(ctx, edge=keys)   # {"path": sorted uint64 array, ...}
    """

    def __init__(self, source, slog=False):
        """source: Datalog text, or with slog=True a Slog compiler's RA plan
        (an s-expression like the reference's test/souffle/cspa.slogc)."""
        _, self.prog = load()
        h = c_void_p()
        fn = self.prog.gdl_datalog_compile_slog if slog else self.prog.gdl_datalog_compile
        st = fn(source.encode(), ctypes.byref(h))
        if st != GDL_OK:
            msg = (self.prog.gdl_datalog_last_error() or b"").decode(errors="replace")
            raise GdlError(st, f"datalog compile: {msg}")
        self.h = h

    @property
    def plan(self):
        return self.prog.gdl_datalog_plan(self.h).decode()

    @property
    def source(self):
        """The Datalog text compiled (for a Slog plan, the rules read back)."""
        return self.prog.gdl_datalog_source(self.h).decode()

    def run(self, ctx, inputs=None, **kw):
        """Runs on ctx with {relation: packed keys} for every .input relation;
        returns {kept relation: sorted duplicate-free uint64 array}."""
        inputs = dict(inputs or {}, **kw)
        names = list(inputs)
        arrs = [np.ascontiguousarray(inputs[n], dtype=np.uint64) for n in names]
        k = len(names)
        c_names = (c_char_p * max(k, 1))(*[n.encode() for n in names])
        c_keys = (_U64P * max(k, 1))(*[_p64(a) for a in arrs])
        c_counts = (c_uint64 * max(k, 1))(*[a.size for a in arrs])
        st = self.prog.gdl_datalog_run(ctx.h, self.h, k, c_names, c_keys, c_counts)
        if st != GDL_OK:
            msg = (self.prog.gdl_datalog_last_error() or b"").decode(errors="replace")
            raise GdlError(st, f"datalog run: {msg}")
        return {n: self.result(n) for n in self.kept()}

    def kept(self):
        """Relations whose tuples run() keeps: the .output / .printsize ones,
        else every derived one."""
        out, i = [], 0
        while True:
            n = self.prog.gdl_datalog_kept(self.h, i)
            if n is None:
                return out
            out.append(n.decode())
            i += 1

    def arity(self, relation):
        """Declared arity (1 or 2); 0 for an unknown relation.  A unary
        relation's tuples are keys c0 << 32."""
        return self.prog.gdl_datalog_arity(self.h, relation.encode())

    def result(self, relation):
        n = c_uint64()
        st = self.prog.gdl_datalog_result(self.h, relation.encode(), None, 0, ctypes.byref(n))
        if st != GDL_OK:
            msg = (self.prog.gdl_datalog_last_error() or b"").decode(errors="replace")
            raise GdlError(st, msg)
        out = np.empty(max(n.value, 1), dtype=np.uint64)
        st = self.prog.gdl_datalog_result(self.h, relation.encode(), _p64(out), n.value,
                                          ctypes.byref(n))
        if st != GDL_OK:
            raise GdlError(st, f"datalog result {relation}")
        return out[:n.value]

    def stats(self):
        """(new tuples, [iterations of each LIE]) of the last run."""
        d = c_uint64()
        n = c_uint32()
        self.prog.gdl_datalog_stats(self.h, ctypes.byref(d), None, 0, ctypes.byref(n))
        its = (c_int32 * max(n.value, 1))()
        self.prog.gdl_datalog_stats(self.h, ctypes.byref(d), its, n.value, ctypes.byref(n))
        return d.value, list(its[:n.value])

    def close(self):
        if getattr(self, "h", None):
            self.prog.gdl_datalog_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fmix64(keys):
    """MurmurHash3 fmix64 of u64 keys (numpy, wrapping)."""
    k = np.asarray(keys, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        k ^= k >> np.uint64(33)
        k *= np.uint64(0xff51afd7ed558ccd)
        k ^= k >> np.uint64(33)
        k *= np.uint64(0xc4ceb9fe1a85ec53)
        k ^= k >> np.uint64(33)
    return k


def tuple_checksum(keys):
    """Order-independent checksum of a tuple set: sum of fmix64, mod 2^64."""
    with np.errstate(over="ignore"):
        return int(fmix64(keys).sum(dtype=np.uint64))


def tc_estimate(edges, n_nodes, samples=32, seed=0):
    """Estimated |TC| (breadth-first search from sampled sources; exact when
    samples >= n_nodes).  Host code of libgdl_hip.so; no device needed."""
    eng, _ = load()
    e = np.ascontiguousarray(edges, dtype=np.uint64)
    return int(eng.gdl_tc_estimate(e.ctypes.data, e.size, n_nodes, samples, seed))


class TcBatch:
    """Many TC fixpoints in one disjoint union (include/gdl_batch.h)."""

    def __init__(self, ctx, max_slots=256, node_bits=16, mem_budget=0, keep_keys=True,
                 solo_tuples=None, solo_threads=None):
        """solo_tuples: instances estimated at this many tuples or more run alone at
        admission (None: the library default; 0: all; ~0: none)."""
        self.ctx = ctx
        cfg = TcbConfig()
        ctx.eng.gdl_tcb_config_default(ctypes.byref(cfg))
        cfg.max_slots, cfg.node_bits, cfg.mem_budget = max_slots, node_bits, mem_budget
        cfg.keep_keys = int(keep_keys)
        if solo_tuples is not None:
            cfg.solo_tuples = int(solo_tuples) & 0xffffffffffffffff
        if solo_threads is not None:
            cfg.solo_threads = int(solo_threads)
        h = c_void_p()
        ctx._check(ctx.eng.gdl_tcb_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h)),
                   "tcb_create")
        self.h = h

    def close(self):
        if self.h:
            self.ctx.eng.gdl_tcb_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def admit(self, qid, edges, n_nodes, est=0):
        """True if admitted, False if the batch is busy (no slot or budget)."""
        e = np.ascontiguousarray(edges, dtype=np.uint64)
        st = self.ctx.eng.gdl_tcb_admit(self.h, qid, e.ctypes.data, e.size, n_nodes, est)
        if st == GDL_ERR_BUSY:
            return False
        self.ctx._check(st, "tcb_admit")
        return True

    def step(self):
        n = c_uint32()
        self.ctx._check(self.ctx.eng.gdl_tcb_step(self.h, ctypes.byref(n)), "tcb_step")
        return n.value

    def poll(self):
        """Finished results (dicts) in completion order, fetched and dropped."""
        out = []
        r = TcbResult()
        while self.ctx.eng.gdl_tcb_poll(self.h, ctypes.byref(r)):
            keys = np.zeros(r.n_out if r.has_keys else 0, dtype=np.uint64)
            tr = (TraceRec * max(int(r.n_trace), 1))()
            self.ctx._check(self.ctx.eng.gdl_tcb_fetch(
                self.h, r.qid, keys.ctypes.data if keys.size else None, tr), "tcb_fetch")
            out.append(dict(qid=r.qid, n_out=r.n_out, checksum=r.checksum, derived=r.derived,
                            iterations=r.iterations, slot=r.slot, admit_ms=r.admit_ms,
                            done_ms=r.done_ms, keys=keys if r.has_keys else None,
                            trace=[(tr[i].iteration, tr[i].relation, tr[i].newt, tr[i].delta,
                                    tr[i].full) for i in range(r.n_trace)]))
        return out

    def info(self):
        v = [c_uint64() for _ in range(5)]
        self.ctx._check(self.ctx.eng.gdl_tcb_info(self.h, *[ctypes.byref(x) for x in v]),
                        "tcb_info")
        return dict(zip(("bodies", "admitted", "finished", "reserved", "budget"),
                        [x.value for x in v]))

    def run(self, instances, est=None):
        """Runs (qid, edges, n_nodes) instances to completion, admitting in
        order as slots and budget allow; returns results by qid."""
        pending = list(instances)
        done = {}
        while pending or len(done) < len(instances):
            while pending:
                qid, edges, n_nodes = pending[0]
                if not self.admit(qid, edges, n_nodes, 0 if est is None else est.get(qid, 0)):
                    break
                pending.pop(0)
            self.step()
            for r in self.poll():
                done[r["qid"]] = r
        return done


# ------------------------------------------------------------------ serving

QUERY_TC, QUERY_SG, QUERY_CSPA = 0, 1, 2


class ServerConfig(Structure):
    _fields_ = [("device", c_int), ("workers", c_int), ("batch_tc", c_int),
                ("batch_slots", c_uint32), ("node_bits", c_uint32), ("mem_budget", c_uint64),
                ("keep_keys", c_int), ("partition_cus", c_int), ("take_turns", c_int)]


class QueryResult(Structure):
    _fields_ = [("qid", c_uint64), ("kind", c_int32), ("status", c_int32), ("n_out", c_uint64),
                ("checksum", c_uint64), ("derived", c_uint64), ("iterations", c_int32),
                ("init_iterations", c_int32), ("est_bytes", c_uint64), ("submit_ms", c_double),
                ("admit_ms", c_double), ("done_ms", c_double), ("batched", c_int32),
                ("rerun_alone", c_int32), ("n_trace", c_uint64), ("n_keys", c_uint64)]


def _serve_lib():
    _, prog = load()
    if not hasattr(prog, "_serve_declared"):
        vp = c_void_p
        _decl(prog, "gdl_server_config_default", None, [POINTER(ServerConfig)])
        _decl(prog, "gdl_server_create", c_int, [POINTER(ServerConfig), POINTER(vp)])
        _decl(prog, "gdl_server_destroy", c_int, [vp])
        _decl(prog, "gdl_query_estimate", c_uint64, [c_int, vp, c_uint64, vp, c_uint64])
        _decl(prog, "gdl_server_submit", c_int, [vp, c_uint64, c_int, vp, c_uint64, vp, c_uint64])
        _decl(prog, "gdl_server_poll", c_int, [vp, POINTER(QueryResult)])
        _decl(prog, "gdl_server_wait", c_int, [vp, POINTER(QueryResult), c_double])
        _decl(prog, "gdl_server_fetch", c_int, [vp, c_uint64, vp, vp])
        _decl(prog, "gdl_server_stats", c_int, [vp] + [_U64P] * 7)
        prog._serve_declared = True
    return prog


def query_estimate(kind, a, b=None):
    """Admission estimate (device bytes) of a query; host code, no device needed."""
    prog = _serve_lib()
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b if b is not None else np.zeros(0), dtype=np.uint64)
    return int(prog.gdl_query_estimate(kind, a.ctypes.data if a.size else None, a.size,
                                       b.ctypes.data if b.size else None, b.size))


class Server:
    """Query server of include/gdl_serve.h: TC / SG / CSPA queries admitted by
    footprint estimate, TC queries continuously batched, the others on worker
    contexts; with take_turns the batch and the workers take turns on the
    device instead of running concurrently."""

    def __init__(self, device=0, workers=1, batch_tc=True, batch_slots=64, node_bits=18,
                 mem_budget=0, keep_keys=False, partition_cus=False, take_turns=False):
        self.prog = _serve_lib()
        cfg = ServerConfig()
        self.prog.gdl_server_config_default(ctypes.byref(cfg))
        cfg.device, cfg.workers, cfg.batch_tc = device, workers, int(batch_tc)
        cfg.batch_slots, cfg.node_bits, cfg.mem_budget = batch_slots, node_bits, mem_budget
        cfg.keep_keys, cfg.partition_cus = int(keep_keys), int(partition_cus)
        cfg.take_turns = int(take_turns)
        h = c_void_p()
        st = self.prog.gdl_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != GDL_OK:
            raise GdlError(st, "gdl_server_create")
        self.h = h

    def close(self):
        if self.h:
            self.prog.gdl_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, qid, kind, a, b=None):
        a = np.ascontiguousarray(a, dtype=np.uint64)
        b = np.ascontiguousarray(b if b is not None else np.zeros(0), dtype=np.uint64)
        st = self.prog.gdl_server_submit(self.h, qid, kind, a.ctypes.data if a.size else None,
                                         a.size, b.ctypes.data if b.size else None, b.size)
        if st != GDL_OK:
            raise GdlError(st, "gdl_server_submit")

    def wait(self, timeout_ms=-1.0):
        """Next finished query as a dict (keys / trace fetched), or None on timeout."""
        r = QueryResult()
        if not self.prog.gdl_server_wait(self.h, ctypes.byref(r), float(timeout_ms)):
            return None
        keys = np.zeros(r.n_keys, dtype=np.uint64)
        tr = (TraceRec * max(int(r.n_trace), 1))()
        st = self.prog.gdl_server_fetch(self.h, r.qid, keys.ctypes.data if keys.size else None,
                                        tr)
        if st != GDL_OK:
            raise GdlError(st, "gdl_server_fetch")
        return dict(qid=r.qid, kind=r.kind, status=r.status, n_out=r.n_out,
                    checksum=r.checksum, derived=r.derived, iterations=r.iterations,
                    init_iterations=r.init_iterations, est_bytes=r.est_bytes,
                    submit_ms=r.submit_ms, admit_ms=r.admit_ms, done_ms=r.done_ms,
                    batched=bool(r.batched), rerun_alone=bool(r.rerun_alone),
                    keys=keys if r.n_keys else None,
                    trace=[(tr[i].iteration, tr[i].relation, tr[i].newt, tr[i].delta,
                            tr[i].full) for i in range(r.n_trace)])

    def poll(self):
        return self.wait(0.0)

    def stats(self):
        v = [c_uint64() for _ in range(7)]
        st = self.prog.gdl_server_stats(self.h, *[ctypes.byref(x) for x in v])
        if st != GDL_OK:
            raise GdlError(st, "gdl_server_stats")
        return dict(zip(("submitted", "finished", "batched", "reruns", "reserved",
                         "peak_reserved", "budget"), [x.value for x in v]))
