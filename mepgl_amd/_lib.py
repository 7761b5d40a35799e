"""ctypes bindings of libmepgl.so (include/mepgl.h).

The HIP library is the only compute path: if it is missing this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmepgl.so")

OK, ERR_ARG, ERR_HIP, ERR_NONFINITE, ERR_STATE, ERR_NOMEM = 0, -1, -2, -3, -4, -5
FP64, MIXED = 0, 1
POLY_EXACT, CUBIC_TEXTBOOK, NO_GRAPH, GAUGE_ALIGN, ROLLBACK, GAUGE_ALIGN_LOCAL = 1, 2, 4, 8, 16, 32

SCALARS = {
    "fenergy": 0, "g_norm2": 1, "alpha": 2, "s": 3, "s_jb": 4, "beta": 5, "g_delta_g": 6,
    "poly": 7, "energy": 8, "g_norm2_old": 9, "ls_index": 10, "theta": 11,
    "winding_1": 12, "winding_2": 13,
}
CONTROL_STOP, CONTROL_RELAXATION_STEPS, CONTROL_NLCG, CONTROL_FOLD_LINE_SEARCH = 0, 1, 2, 3
CONTROL_SPLIT_IMAGES = 4
CONTROL_EDGE_EXCHANGE = 5
KERNEL_GRAD, KERNEL_POLY = 0, 1

EXPORTS = [
    "mepgl_create", "mepgl_destroy", "mepgl_last_error", "mepgl_upload", "mepgl_download",
    "mepgl_reset_solver", "mepgl_free_energy", "mepgl_gradient", "mepgl_nlcg_steps",
    "mepgl_compute_jb", "mepgl_arc_length", "mepgl_redistribute", "mepgl_resample",
    "mepgl_reset_interior_beta", "mepgl_gauge_align", "mepgl_gauge_align_local", "mepgl_loop", "mepgl_get_scalars", "mepgl_download_state",
    "mepgl_set_control", "mepgl_get_control", "mepgl_profile", "mepgl_profile_read",
    "mepgl_synchronize", "mepgl_info", "mepgl_set_segment_rows", "mepgl_autotune", "mepgl_upload_state", "mepgl_set_scalars", "mepgl_winding",
    "mepgl_free_energy_noalloc",
]

_dp = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


class Param(C.Structure):
    _fields_ = [("value", C.c_double), ("plane", _dp)]


class StringDesc(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("F", C.c_int), ("ncomp", C.c_int), ("no_self_field", C.c_int),
        ("dx", C.c_double),
        ("a_1", Param), ("b_1", Param), ("m_xx_1", Param), ("m_yy_1", Param),
        ("a_2", Param), ("b_2", Param), ("m_xx_2", Param), ("m_yy_2", Param),
        ("h", Param),
        ("q_1", C.c_double), ("q_2", C.c_double), ("eta", C.c_double), ("gamma", C.c_double),
        ("delta", C.c_double),
        ("comp_domain", _i32p), ("sc_domain", _i32p),
        ("segment_rows", C.c_int),
    ]


class MepglError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmepgl error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libmepgl.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mepgl_amd.build` "
                          "(there is no non-HIP fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.mepgl_create.argtypes = [C.POINTER(StringDesc), C.c_int, C.c_int, C.c_uint, C.c_int,
                               C.POINTER(vp)]
    L.mepgl_destroy.argtypes = [vp]
    L.mepgl_destroy.restype = None
    L.mepgl_last_error.argtypes = [vp]
    L.mepgl_last_error.restype = C.c_char_p
    L.mepgl_upload.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.mepgl_download.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.mepgl_reset_solver.argtypes = [vp]
    L.mepgl_free_energy.argtypes = [vp]
    L.mepgl_winding.argtypes = [vp]
    L.mepgl_gradient.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.mepgl_nlcg_steps.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.mepgl_compute_jb.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    L.mepgl_arc_length.argtypes = [vp]
    L.mepgl_redistribute.argtypes = [vp, C.c_int]
    L.mepgl_resample.argtypes = [vp, C.c_int, C.c_int]
    L.mepgl_reset_interior_beta.argtypes = [vp]
    L.mepgl_gauge_align.argtypes = [vp]
    L.mepgl_gauge_align_local.argtypes = [vp]
    L.mepgl_loop.argtypes = [vp, C.c_int, C.c_int, C.c_char, C.c_int]
    L.mepgl_get_scalars.argtypes = [vp, C.c_int, C.c_int, _dp]
    L.mepgl_download_state.argtypes = [vp, C.c_int, C.c_int, _dp, _dp, _dp]
    L.mepgl_upload_state.argtypes = [vp, C.c_int, C.c_int, _dp, _dp, _dp]
    L.mepgl_set_scalars.argtypes = [vp, C.c_int, C.c_int, _dp]
    L.mepgl_set_control.argtypes = [vp, C.c_int, C.c_int]
    L.mepgl_get_control.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.mepgl_profile.argtypes = [vp, C.c_int]
    L.mepgl_profile_read.argtypes = [vp, C.c_int, _dp, C.POINTER(C.c_int)]
    L.mepgl_synchronize.argtypes = [vp]
    L.mepgl_set_segment_rows.argtypes = [vp, C.c_int, C.c_int]
    L.mepgl_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.mepgl_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_longlong)]
    _lib = L
    return L
