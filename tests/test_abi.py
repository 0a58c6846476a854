"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import pf_binding as pfb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pf_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"pf_plan_create", "pf_generate", "pf_matvec", "pf_cg", "pf_eigs"} <= names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pfb.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_names_are_exported():
    lib = ctypes.CDLL(pfb.LIB_PATH)
    for n in pfb.exported_symbols():
        assert hasattr(lib, n), n


def test_desc_struct_layout_matches_header():
    # offsets of the C struct as compiled by the C compiler
    assert ctypes.sizeof(pfb.PfDesc) == 224
    assert ctypes.sizeof(pfb.PfEigStats) == 40
    assert ctypes.sizeof(pfb.PfSizes) == 64
