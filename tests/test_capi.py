"""The C-ABI library loads and exports every entry point include/magreete_hip.h
declares, and the ctypes table matches the header (no GPU calls)."""
import os
import re

import pytest

from magreete_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mg_\w+)\s*\(", src, re.M))
    return names


def test_header_declares_entry_points():
    names = declared()
    for n in ("mg_assemble", "mg_getrf", "mg_getrs", "mg_propagate", "mg_cgemm", "mg_coldot", "mg_sources"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.mg_version() >= 1


def test_ctypes_table_covers_header():
    assert declared() == set(_lib.SIGNATURES)


def _header_params():
    """name -> list of parameter C types, from include/*.h."""
    out = {}
    for f in os.listdir(os.path.join(ROOT, "include")):
        if not f.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", f)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"(mg_\w+)\s*\(([^)]*)\)\s*;", src):
            params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
            out[m.group(1)] = [re.sub(r"\s*\b\w+$", "", p) if not p.endswith("*") else p for p in params]
    return out


def _ctype_of(decl):
    import ctypes
    d = decl.replace("const ", "").strip()
    if d.endswith("**"):
        return ctypes.POINTER(ctypes.c_void_p)
    if d.endswith("*"):
        base = d[:-1].strip()
        return {"long long": ctypes.POINTER(ctypes.c_longlong),
                "double": ctypes.POINTER(ctypes.c_double)}.get(base, ctypes.c_void_p)
    return {"int": ctypes.c_int, "int64_t": ctypes.c_int64, "double": ctypes.c_double}[d]


def test_ctypes_argument_types_match_header():
    hp = _header_params()
    for name, args in _lib.SIGNATURES.items():
        want = [_ctype_of(p) for p in hp[name]]
        assert len(want) == len(args), name
        for i, (a, b) in enumerate(zip(args, want)):
            ptr_a = a.__name__ == "c_void_p" or a.__name__.startswith("LP_")
            ptr_b = b.__name__ == "c_void_p" or b.__name__.startswith("LP_")
            if ptr_a and ptr_b:
                continue   # device buffers are bound as void*, host out-pointers as typed pointers
            assert a is b, (name, i, a, b)
