"""The C-ABI library loads and exports every entry point include/tpx.h
declares (no device calls)."""
import os
import re

import tpx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tpx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tpx_\w+)\s*\(", text, re.M)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(tpx.EXPORTS)


def test_library_exports_all_symbols():
    L = tpx.load()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_struct_layout_matches_header():
    text = open(os.path.join(ROOT, "include", "tpx.h")).read()
    body = text[text.index("typedef struct tpx_params"):text.index("} tpx_params;")]
    names = re.findall(r"\b(\w+)\s*[,;]", body.split("{", 1)[1])
    assert names == [n for n, _ in tpx.TpxParams._fields_]


def test_create_rejects_bad_sizes_without_device():
    import ctypes
    p = tpx.Params(Nx_tot=24)
    st = tpx.TpxParams(**p.struct_fields())
    h = ctypes.c_void_p()
    rc = tpx.load().tpx_create(ctypes.byref(st), 0, ctypes.byref(h))
    assert rc == tpx.TPX_E_ARG
    assert b"power" in tpx.load().tpx_create_error()
