"""The C ABI library loads on a CPU-only host and exports every symbol include/gvdbx.h declares."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gvdbx.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gvdbx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_path():
    syms = declared_symbols()
    for s in ("gvdbx_rebuild_topology", "gvdbx_update_atlas", "gvdbx_insert_points_subcell", "gvdbx_gather_density",
              "gvdbx_gather_levelset", "gvdbx_scatter_reduce_levelset", "gvdbx_update_apron", "gvdbx_configure"):
        assert s in syms


def test_library_exports_all_declared_symbols():
    from gvdbx.lib import LIB_PATH, SIGNATURES
    if not os.path.exists(LIB_PATH):
        pytest.fail("libgvdbx.so is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(SIGNATURES), "ctypes table and header disagree"


def test_status_strings_and_cpu_only_context():
    from gvdbx.lib import load_library
    L = load_library()
    assert L.gvdbx_status_string(0) == b"ok"
    assert L.gvdbx_status_string(3).startswith(b"point extents")
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and not os.path.exists("/dev/kfd"):
        h = C.c_void_p()
        assert L.gvdbx_ctx_create(0, None, C.byref(h)) == 5  # GVDBX_ERR_HIP, no crash
        assert not h.value
    assert L.gvdbx_configure(None, 5, None) == 1
    assert L.gvdbx_rebuild_topology(None, 10, 1.0, None) == 1
