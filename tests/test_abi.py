"""The C-ABI library loads without a GPU and exports every entry point the
public header declares (no compute call is made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "broadcast_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(bca_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "bca_residual_dnc5" in names and "bca_jvp_dnc5" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from broadcast_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libbroadcast_amd.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(declared()) == sorted(_lib.EXPORTS)


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "broadcast_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", src, re.M), f
