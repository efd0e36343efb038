"""The C ABI libraries load and export every symbol the headers declare (no GPU
compute here: only symbol resolution and the status-string table)."""
import ctypes
import os
import re

import pytest

from conftest import REPO, gdl

HEADERS = {
    "libgdl_hip.so": [os.path.join(REPO, "include", "gdl_hip.h"),
                      os.path.join(REPO, "include", "gdl_batch.h")],
    "libgdlog.so": [os.path.join(REPO, "include", "gdl_programs.h"),
                    os.path.join(REPO, "include", "gdl_serve.h"),
                    os.path.join(REPO, "include", "gdl_datalog.h")],
}


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gdl_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("lib", sorted(HEADERS))
def test_library_exports_every_declared_symbol(lib):
    names = sorted({n for h in HEADERS[lib] for n in declared(h)})
    assert names, "no declarations parsed"
    L = ctypes.CDLL(gdl.lib_path(lib), mode=ctypes.RTLD_GLOBAL)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{lib} lacks {missing}"


def test_status_strings_and_defaults():
    eng, _ = gdl.load()
    assert eng.gdl_status_string(0) == b"ok"
    assert eng.gdl_status_string(2) == b"out of device memory"
    cfg = gdl.Config()
    eng.gdl_config_default(ctypes.byref(cfg))
    assert cfg.device == 0 and abs(cfg.load_factor - 0.35) < 1e-6 and cfg.mem_cap_bytes == 0


def test_null_arguments_are_rejected_without_a_device():
    eng, prog = gdl.load()
    # every entry validates its context pointer before touching HIP
    assert eng.gdl_sync(None) == 1
    assert eng.gdl_alloc(None, 16, None) == 1
    assert eng.gdl_sort_u64(None, None, 0, 32, 32, None) == 1
    res = gdl.ProgResult()
    assert prog.gdl_prog_tc(None, None, 0, 0, -1, ctypes.byref(res), None) == 1
    assert prog.gdl_prog_path_length(None, None, 0, 0, ctypes.byref(res), None, None, 0) == 1
    assert eng.gdl_segment_labels_u32(None, None, 1, 0, None, 0) == 1
    assert eng.gdl_segment_labels_sorted(None, None, 0, None, 1, 0, 8, None, None) == 1
    assert eng.gdl_graph_begin(None) == 1
    assert eng.gdl_graph_end(None, None) == 1
    assert eng.gdl_graph_launch(None, None) == 1
    assert eng.gdl_graph_destroy(None) == 0  # destroying nothing is fine
    # the Datalog front end compiles without a device; running needs a context
    p = ctypes.c_void_p()
    assert prog.gdl_datalog_compile(None, ctypes.byref(p)) == 1
    assert prog.gdl_datalog_compile(b".decl r(a: number, b: number)", ctypes.byref(p)) == 0
    try:
        assert prog.gdl_datalog_run(None, p, 0, None, None, None) == 1
        n = ctypes.c_uint64()
        assert prog.gdl_datalog_result(p, b"r", None, 0, ctypes.byref(n)) == 1  # not run yet
        assert prog.gdl_datalog_kept(p, 0) == b"r" and prog.gdl_datalog_kept(p, 1) is None
        assert prog.gdl_datalog_arity(p, b"r") == 2 and prog.gdl_datalog_arity(p, b"x") == 0
    finally:
        prog.gdl_datalog_destroy(p)


def test_generator_library_is_deterministic():
    a = gdl.gen_er(0, 2000, 5000)
    b = gdl.gen_er(0, 2000, 5000)
    assert a.size == 5000 and (a == b).all()
    s, d = gdl.unpack(a)
    assert (s != d).all() and len(set(a.tolist())) == 5000
    assert s.max() < 2000 and d.max() < 2000
    t = gdl.gen_tree(0, 200000, 1000)
    c, p = gdl.unpack(t)
    assert t.size == 199999 and (p < c).all() and (c - p <= 1000).all()
    r = gdl.gen_rmat(1, 15, 40000)
    rs, rd = gdl.unpack(r)
    assert r.size == 40000 and len(set(r.tolist())) == 40000 and (rs != rd).all()
    assert rs.max() < 32768 and rd.max() < 32768


def test_cu_masks_are_refused_without_a_device():
    eng, _ = gdl.load()
    prog = gdl._serve_lib()
    assert eng.gdl_set_cu_mask(None, None, 0) == 1  # null context
    # a server asked for per-executor CU masks is refused before any device call
    cfg = gdl.ServerConfig()
    prog.gdl_server_config_default(ctypes.byref(cfg))
    cfg.partition_cus = 1
    h = ctypes.c_void_p()
    assert prog.gdl_server_create(ctypes.byref(cfg), ctypes.byref(h)) == 4
    assert not h.value
