"""The C ABI library loads and exports every function include/upc_lumi.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "upc_lumi.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(upc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("upc_ctx_create", "upc_lumi_grid", "upc_nuc_cs_ym", "upc_pdf_init", "upc_flux_form"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import upcgen_amd as U
    assert os.path.exists(U.LIB_PATH), "build the library first: python -m upcgen_amd.build"
    lib = ctypes.CDLL(U.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared ABI
    assert sorted(U.SIGNATURES) == declared()


def _header_fields(struct):
    text = open(HEADER).read()
    body = text[text.index(f"typedef struct {struct}"):text.index(f"}} {struct};")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(int|double|long long)\s+([^;]+);", body)
    names = []
    for _, decl in fields:
        names += [d.strip() for d in decl.split(",")]
    return names


def test_config_struct_matches_header():
    import upcgen_amd as U
    assert _header_fields("upc_config") == [f for f, _ in U.Config._fields_]


def test_info_struct_matches_header():
    import upcgen_amd as U
    assert _header_fields("upc_info") == [f for f, _ in U.Info._fields_]


def test_error_path_without_device():
    import upcgen_amd as U
    lib = U.load()
    cfg = U.Config.make(Z=0)
    ctx = ctypes.c_void_p()
    rc = lib.upc_ctx_create(ctypes.byref(cfg), None, ctypes.byref(ctx))
    assert rc == -1 and b"Z" in lib.upc_last_error()
    assert not ctx.value
