"""The C-ABI library loads and exports every symbol include/jhip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from jenhine_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(jh_\w+)\s*\(", txt, re.M))
    return names


def test_header_declarations_are_bound():
    decl = declared_symbols()
    assert len(decl) >= 20
    assert decl == set(_lib.EXPORTS), decl ^ set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_error_path_without_device():
    L = _lib.lib()
    assert L.jh_version() >= 100
    # an invalid call fails with a message and no device work
    h = ctypes.c_void_p()
    st = L.jh_bp_create(ctypes.byref(h), 0, None, 0, 0, 0, 0, 0, 0, ctypes.c_float(0), ctypes.c_float(0))
    assert st == _lib.JH_ERR_INVALID
    assert b"projection" in L.jh_last_error()
