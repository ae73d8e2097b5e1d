"""The C-ABI library loads and exports every entry point declared in
include/diablo_hip.h (no compute call: runs without a GPU)."""
import ctypes
import os
import re

from diablo_amd import channel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "diablo_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "dh_rk_substep" in syms and "dh_create" in syms
    # the ctypes table of the host mirror covers the whole header
    assert sorted(channel.EXPORTS) == syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(channel.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_params_struct_layout_matches_oracle():
    from oracle import dor
    assert ctypes.sizeof(channel.DhParams) == ctypes.sizeof(dor.DorParams)
    for (n1, t1), (n2, t2) in zip(channel.DhParams._fields_, dor.DorParams._fields_):
        assert n1 == n2
        assert ctypes.sizeof(t1) == ctypes.sizeof(t2)
