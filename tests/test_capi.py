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
