"""ctypes binding of libjhip.so (the C-ABI declared in include/jhip.h).

This is the binding a Python host uses; the C++ agents link the same library
directly.  Loading fails loudly when the library is missing: there is no
fallback implementation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjhip.so")

JH_OK, JH_ERR_INVALID, JH_ERR_OOM, JH_ERR_HIP = 0, 1, 2, 3
JH_F32, JH_BF16, JH_F16 = 0, 1, 2
JH_FILTER_FP32, JH_FILTER_MFMA_BF16X3, JH_FILTER_MFMA_BF16X6, JH_FILTER_FP32_VALU, JH_FILTER_MFMA_F16X3 = 0, 1, 2, 3, 4
JH_WINDOW_REFERENCE, JH_WINDOW_SHEPP_LOGAN = 0, 1
JH_LAYOUT_QUAD, JH_LAYOUT_STRIP = 0, 1
JH_WEIGHT_COUNT, JH_WEIGHT_FDK = 0, 1
JH_NOALLOC_FIRST, JH_NOALLOC_LAST = 1, 2
JH_PTM_MOVE_U, JH_PTM_MOVE_V, JH_PTM_REVERSE_X, JH_PTM_REVERSE_Y, JH_PTM_MOVE_Z = range(5)

DTYPES = {"f32": JH_F32, "bf16": JH_BF16, "f16": JH_F16}

# Every symbol include/jhip.h declares (tests check the library exports all).
EXPORTS = [
    "jh_last_error", "jh_version", "jh_device_count", "jh_ptm_modify",
    "jh_pre_create", "jh_pre_run", "jh_pre_output_bound", "jh_pre_destroy",
    "jh_filter_create", "jh_filter_create_ex", "jh_filter_shape", "jh_filter_response", "jh_filter_preweight",
    "jh_filter_run", "jh_filter_stage_fusable", "jh_filter_run_staged", "jh_filter_operand_bytes",
    "jh_pre_run_operand", "jh_filter_run_operand_staged", "jh_filter_output_bound", "jh_filter_set_generated", "jh_filter_get_generated", "jh_filter_destroy",
    "jh_detector_on_right", "jh_composite_create", "jh_composite_side", "jh_composite_run", "jh_composite_run_operand",
    "jh_composite_destroy",
    "jh_bp_create", "jh_bp_uweights", "jh_bp_stage_bytes", "jh_bp_stage", "jh_bp_run", "jh_bp_reset",
    "jh_bp_accumulate", "jh_bp_accumulate_slab", "jh_bp_finalize", "jh_bp_accumulate_noalloc", "jh_bp_set_tile", "jh_bp_get_tile", "jh_bp_geometry_mode",
    "jh_bp_destroy", "jh_bp_pack", "jh_bp_backproject", "jh_fp_project", "jh_bp_autotune",
    "jh_bp_set_fetch", "jh_bp_get_fetch", "jh_bp_fetch_in_use", "jh_bp_set_layout", "jh_bp_get_layout", "jh_bp_set_zfilter",
    "jh_bp_set_ctnum", "jh_bp_set_weight", "jh_bp_get_weight", "jh_bp_set_f16_exponent", "jh_bp_get_f16_exponent",
    "jh_zfilter_run",
    "jh_ctnum_run",
    "jh_graph_create", "jh_graph_add_scan", "jh_graph_autotune", "jh_graph_scan_tuning", "jh_graph_instantiate",
    "jh_graph_launch", "jh_graph_destroy",
]


class JhError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"jhip error {status}: {msg}")
        self.status = status


class JhOutOfMemory(JhError):
    pass


_lib = None

c_int, c_float, c_double, c_void_p, c_size_t = (ctypes.c_int, ctypes.c_float, ctypes.c_double,
                                               ctypes.c_void_p, ctypes.c_size_t)
P = c_void_p

_SIGS = {
    "jh_last_error": (ctypes.c_char_p, []),
    "jh_version": (c_int, []),
    "jh_device_count": (c_int, []),
    "jh_ptm_modify": (c_int, [P, c_int, c_int, c_double, P]),
    "jh_pre_create": (c_int, [P, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_float, P]),
    "jh_pre_run": (c_int, [P, P, P, P, P, c_int, P]),
    "jh_pre_destroy": (c_int, [P]),
    "jh_filter_create": (c_int, [P, c_int, c_int, c_int, c_int, P, c_float, c_float, c_float, c_float, P, P,
                                 c_int, c_int, c_int]),
    "jh_filter_create_ex": (c_int, [P, c_int, c_int, c_int, c_int, c_int, c_int, P, c_float, c_float, c_float,
                                    c_float, P, P, c_int, c_int, c_int]),
    "jh_filter_shape": (c_int, [P, P, P, P]),
    "jh_filter_response": (c_int, [P, P, P]),
    "jh_filter_preweight": (c_int, [P, c_int, P, P]),
    "jh_filter_run": (c_int, [P, P, P, c_int, c_int, P]),
    "jh_filter_stage_fusable": (c_int, [P, P, c_int, P]),
    "jh_filter_run_staged": (c_int, [P, P, P, c_int, c_int, c_int, P, P]),
    "jh_filter_operand_bytes": (c_int, [P, c_int, P]),
    "jh_pre_run_operand": (c_int, [P, P, P, P, P, c_int, c_int, P, P]),
    "jh_filter_run_operand_staged": (c_int, [P, P, P, c_int, c_int, c_int, P, P]),
    "jh_filter_set_generated": (c_int, [P, c_int]),
    "jh_filter_get_generated": (c_int, [P, P]),
    "jh_filter_destroy": (c_int, [P]),
    "jh_detector_on_right": (c_int, [P, c_int]),
    "jh_composite_create": (c_int, [P, c_int, c_int, c_int, c_int, P, c_float, c_int]),
    "jh_composite_side": (c_int, [P, P]),
    "jh_composite_run": (c_int, [P, P, P, ctypes.c_longlong, P, c_int, c_int, P]),
    "jh_composite_run_operand": (c_int, [P, P, P, P, ctypes.c_longlong, c_int, c_int, P, c_int, P]),
    "jh_composite_destroy": (c_int, [P]),
    "jh_bp_create": (c_int, [P, c_int, P, c_int, c_int, c_int, c_int, c_int, c_int, c_float, c_float]),
    "jh_bp_uweights": (c_int, [P, P]),
    "jh_bp_stage_bytes": (c_int, [P, c_int, c_int, P]),
    "jh_bp_stage": (c_int, [P, P, c_int, c_int, c_int, P, P]),
    "jh_bp_run": (c_int, [P, P, c_int, P, P]),
    "jh_bp_reset": (c_int, [P, P]),
    "jh_bp_accumulate": (c_int, [P, P, c_int, c_int, c_int, P]),
    "jh_bp_accumulate_slab": (c_int, [P, P, c_int, c_int, c_int, c_int, c_int, P]),
    "jh_bp_finalize": (c_int, [P, P, c_int, c_int, P]),
    "jh_bp_set_tile": (c_int, [P, c_int, c_int]),
    "jh_bp_get_tile": (c_int, [P, P, P]),
    "jh_bp_geometry_mode": (c_int, [P, P]),
    "jh_bp_destroy": (c_int, [P]),
    "jh_bp_pack": (c_int, [P, P, c_int, c_int, c_int, P, P]),
    "jh_bp_backproject": (c_int, [P, P, c_int, c_int, P, P]),
    "jh_fp_project": (c_int, [P, P, P, c_int, c_int, c_int, P]),
    "jh_bp_autotune": (c_int, [P, P, c_int, c_int, P, c_int, P, P]),
    "jh_bp_set_fetch": (c_int, [P, c_int]),
    "jh_bp_set_zfilter": (c_int, [P, c_int]),
    "jh_bp_set_ctnum": (c_int, [P, c_int, c_float, c_float, c_float]),
    "jh_bp_set_weight": (c_int, [P, c_int, c_float]),
    "jh_bp_set_f16_exponent": (c_int, [P, c_int]),
    "jh_bp_accumulate_noalloc": (c_int, [P, P, c_int, c_int, c_int, c_int, P, P]),
    "jh_bp_get_f16_exponent": (c_int, [P, P]),
    "jh_pre_output_bound": (c_int, [P, c_double, c_double, P]),
    "jh_filter_output_bound": (c_int, [P, c_double, P]),
    "jh_bp_get_weight": (c_int, [P, P, P]),
    "jh_zfilter_run": (c_int, [P, P, c_int, c_int, c_int, P]),
    "jh_ctnum_run": (c_int, [P, c_size_t, c_float, c_float, c_float, P]),
    "jh_bp_get_fetch": (c_int, [P, P]),
    "jh_bp_fetch_in_use": (c_int, [P, c_int, P, P]),
    "jh_bp_set_layout": (c_int, [P, c_int]),
    "jh_bp_get_layout": (c_int, [P, P]),
    "jh_graph_create": (c_int, [P, c_int]),
    "jh_graph_add_scan": (c_int, [P, P, P, P, P, P, P, c_int, c_int, P, P]),
    "jh_graph_autotune": (c_int, [P, c_int, P]),
    "jh_graph_scan_tuning": (c_int, [P, c_int, P, P, P]),
    "jh_graph_instantiate": (c_int, [P, c_int]),
    "jh_graph_launch": (c_int, [P, P]),
    "jh_graph_destroy": (c_int, [P]),
}


def lib():
    """The loaded library; raises if libjhip.so has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m jenhine_amd.build` "
                               "(there is no fallback path)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != JH_OK:
        msg = lib().jh_last_error().decode(errors="replace")
        if status == JH_ERR_OOM:
            raise JhOutOfMemory(status, msg)
        raise JhError(status, msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))
