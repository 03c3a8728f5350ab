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


def test_detector_side_and_argument_checks_without_device():
    import numpy as np
    from jenhine_amd import configs
    from oracle import oracle as o
    L = _lib.lib()
    for shift in (48.0, -48.0, 0.0):
        P = np.ascontiguousarray(configs.OFFSET128.scaled(u_shift=shift).ptms()[0], np.float64)
        got = L.jh_detector_on_right(P.ctypes.data_as(ctypes.c_void_p), 160)
        assert bool(got) == o.detector_on_right(P, 160) == (shift >= 0.0)
    h = ctypes.c_void_p()
    P = np.zeros((15, 12))
    # an odd view count has no conjugate pairs
    st = L.jh_composite_create(ctypes.byref(h), 0, 64, 32, 15, P.ctypes.data_as(ctypes.c_void_p),
                               ctypes.c_float(300.0), -1)
    assert st == _lib.JH_ERR_INVALID and b"even" in L.jh_last_error()
    # yMirroring keeps output columns [Wout, 2 Wout), which must lie inside the FFT length
    st = L.jh_filter_create_ex(ctypes.byref(h), 0, 64, 8, 80, 1, 4, P.ctypes.data_as(ctypes.c_void_p),
                               ctypes.c_float(300.0), ctypes.c_float(1.0), ctypes.c_float(0), ctypes.c_float(0),
                               None, None, 0, 0, 0)
    assert st == _lib.JH_ERR_INVALID and b"FFT length" in L.jh_last_error()


def test_fused_path_entry_points_reject_bad_arguments_without_device():
    L = _lib.lib()
    ok = ctypes.c_int(7)
    # no plans: not fusable, and the runs refuse before any device work
    assert L.jh_filter_stage_fusable(None, None, _lib.JH_BF16, ctypes.byref(ok)) == _lib.JH_OK
    assert ok.value == 0
    assert L.jh_filter_stage_fusable(None, None, _lib.JH_BF16, None) == _lib.JH_ERR_INVALID
    st = L.jh_filter_run_staged(None, None, None, 0, 1, _lib.JH_BF16, None, None)
    assert st == _lib.JH_ERR_INVALID and b"bad arguments" in L.jh_last_error()
    st = L.jh_pre_run_operand(None, None, None, None, None, 0, 1, None, None)
    assert st == _lib.JH_ERR_INVALID and b"bad arguments" in L.jh_last_error()
    st = L.jh_filter_run_operand_staged(None, None, None, 0, 1, _lib.JH_BF16, None, None)
    assert st == _lib.JH_ERR_INVALID and b"bad arguments" in L.jh_last_error()
    st = L.jh_composite_run_operand(None, None, None, None, ctypes.c_longlong(0), 0, 1, None, 0, None)
    assert st == _lib.JH_ERR_INVALID and b"bad arguments" in L.jh_last_error()
    b = ctypes.c_size_t()
    assert L.jh_filter_operand_bytes(None, 1, ctypes.byref(b)) == _lib.JH_ERR_INVALID
