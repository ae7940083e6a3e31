"""ctypes binding of libzorro_hip.so (declared in include/zorro_hip.h).

This is the binding a maintainer would place beside the reference's
pyFFTW/numexprz path.  Loading fails loudly when the library is missing: the
product path has no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzorro_hip.so")

ZR_OK, ZR_ERR_ARG, ZR_ERR_HIP, ZR_ERR_UNSUPPORTED, ZR_ERR_STATE, ZR_ERR_NOMEM = 0, -1, -2, -3, -4, -5
ZR_TRI_TRI, ZR_TRI_DIAG, ZR_TRI_AUTO, ZR_TRI_FIRST, ZR_TRI_REFINE = 0, 1, 2, 3, 4
ZR_B_NONE, ZR_B_CONV, ZR_B_OPTI = 0, 1, 2
ZR_SHIFT_INTEGER, ZR_SHIFT_LANCZOS, ZR_SHIFT_FOURIER = 0, 1, 2
ZR_W_AUTOLOGISTIC, ZR_W_LOGISTIC, ZR_W_CORR, ZR_W_UNWEIGHTED, ZR_W_NORM = 0, 1, 2, 3, 4
ZR_DTYPE_F32, ZR_DTYPE_U8, ZR_DTYPE_U16, ZR_DTYPE_U4, ZR_DTYPE_I16 = 0, 1, 2, 3, 4
ZR_FILTER_DOSE, ZR_FILTER_DOSENORM, ZR_FILTER_GAUSSLP, ZR_FILTER_BACKGROUND = 1, 2, 3, 4


class ZrParams(C.Structure):
    _fields_ = [
        ("tri_mode", C.c_int), ("diag_width", C.c_int), ("diag_start", C.c_int),
        ("max_shift", C.c_int),
        ("b_mode", C.c_int), ("b_rad", C.c_double),
        ("sub_pix_reg", C.c_int),
        ("shift_method", C.c_int),
        ("weight_mode", C.c_int), ("peaksig_thres", C.c_double), ("logistic_k", C.c_double),
        ("logistic_nu", C.c_double),
        ("origin_centroid", C.c_int), ("suppress_origin", C.c_int),
        ("fou_crop", C.c_int * 2), ("shape_padded", C.c_int * 2),
        ("do_dose", C.c_int), ("dose_a", C.c_double), ("dose_b", C.c_double), ("dose_c", C.c_double),
        ("dose_order", C.c_double), ("dose_missing", C.c_double),
        ("pixelsize_nm", C.c_double), ("voltage_kv", C.c_double),
        ("do_frc", C.c_int), ("frc_box", C.c_int), ("frc_overlap", C.c_double),
        ("auto_max", C.c_int),
        ("shape_binned", C.c_int * 2),
        ("do_hot", C.c_int), ("hot_logistic_k", C.c_double), ("hot_relax", C.c_double),
        ("hot_max_sigma", C.c_double), ("hot_psf_k2", C.c_int), ("hot_cutoff_lower", C.c_double),
        ("hot_cutoff_upper", C.c_double), ("hot_fixed_cutoffs", C.c_int), ("hot_neighbour_thres", C.c_double), ("hot_mu_sigma", C.c_double * 2),
        ("pre_shift", C.c_int),
        ("n_filters", C.c_int), ("filters", C.c_int * 4),
        ("k1_batch", C.c_int), ("k2_cols_per_wave", C.c_int), ("k2_pairs_per_wave", C.c_int),
        ("user_mask", C.c_int),
    ]


ZR_REQ_QUEUED, ZR_REQ_RUNNING, ZR_REQ_DONE, ZR_REQ_FAILED, ZR_REQ_CANCELLED = 0, 1, 2, 3, 4


class ZrServeJob(C.Structure):
    """zr_serve_job: one movie as a plan operation of the server sees it."""
    _fields_ = [
        ("params", C.POINTER(ZrParams)), ("frames", C.c_void_p), ("n", C.c_int), ("h", C.c_int), ("w", C.c_int),
        ("image_sum", C.c_void_p), ("filt_sum", C.c_void_p), ("translations", C.POINTER(C.c_double)),
        ("frc", C.POINTER(C.c_double)), ("frc_cap", C.c_int), ("frc_n", C.POINTER(C.c_int)),
    ]


PLAN_ESTIMATE = C.CFUNCTYPE(C.c_size_t, C.c_void_p, C.POINTER(ZrParams), C.c_int, C.c_int, C.c_int)
PLAN_CREATE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(ZrParams), C.c_int, C.c_int, C.c_int,
                          C.POINTER(C.c_void_p), C.POINTER(C.c_size_t))
PLAN_RUN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.POINTER(ZrServeJob))
PLAN_DESTROY = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)


class ZrPlanOps(C.Structure):
    """zr_plan_ops: the plan operations a server drives."""
    _fields_ = [("ctx", C.c_void_p), ("estimate", PLAN_ESTIMATE), ("create", PLAN_CREATE), ("run", PLAN_RUN),
                ("destroy", PLAN_DESTROY)]


class ZrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libzorro_hip error {code}: {msg}")
        self.code = code


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "zr_last_error": (C.c_char_p, []),
    "zr_device_count": (C.c_int, [_I]),
    "zr_movie_create": (C.c_int, [C.c_int, C.POINTER(ZrParams), C.c_int, C.c_int, C.c_int, C.POINTER(_P)]),
    "zr_movie_create_pairs": (C.c_int, [C.c_int, C.POINTER(ZrParams), C.c_int, C.c_int, C.c_int, C.c_void_p,
                                        C.POINTER(_P)]),
    "zr_movie_destroy": (None, [_P]),
    "zr_movie_set_stream": (C.c_int, [_P, _P]),
    "zr_movie_set_frames": (C.c_int, [_P, _P, C.c_int]),
    "zr_movie_set_mask": (C.c_int, [_P, _P, C.c_int]),
    "zr_movie_set_frames_raw": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, C.c_int, C.c_double]),
    "zr_xcorr": (C.c_int, [_P]),
    "zr_solve": (C.c_int, [_P]),
    "zr_set_trimats": (C.c_int, [_P, _F, _F, _F]),
    "zr_apply_shifts": (C.c_int, [_P, _P]),
    "zr_frc": (C.c_int, [_P]),
    "zr_dose_filter": (C.c_int, [_P]),
    "zr_filter": (C.c_int, [_P, C.c_int]),
    "zr_post_filters": (C.c_int, [_P]),
    "zr_get_peaksig_hist": (C.c_int, [_P, _F, _F, _F, C.c_int, _I]),
    "zr_get_background_fit": (C.c_int, [_P, _D, _D, _I]),
    "zr_tiled_frc": (C.c_int, [_P, _P, C.c_int, C.c_int, _D, C.c_int, C.c_int, C.c_double, _P, _I, _P]),
    "zr_tiled_frc_noalloc": (C.c_int, [_P, _P, C.c_int, C.c_int, _D, C.c_int, C.c_int, C.c_double, _P, _I, _P]),
    "zr_batch_create": (C.c_int, [C.c_int, C.POINTER(ZrParams), C.c_int, _I, C.c_int, C.c_int, C.POINTER(_P)]),
    "zr_batch_destroy": (None, [_P]),
    "zr_batch_movie": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "zr_batch_frames": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "zr_batch_set_frames": (C.c_int, [_P, C.c_int, _P, C.c_int]),
    "zr_batch_use_graph": (C.c_int, [_P, C.c_int]),
    "zr_batch_align": (C.c_int, [_P]),
    "zr_batch_synchronize": (C.c_int, [_P]),
    "zr_batch_device_bytes": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "zr_align": (C.c_int, [_P]),
    "zr_synchronize": (C.c_int, [_P]),
    "zr_movie_use_graph": (C.c_int, [_P, C.c_int]),
    "zr_get_trimats": (C.c_int, [_P, _F, _F, _F, _F]),
    "zr_get_windows": (C.c_int, [_P, _F]),
    "zr_get_peak_sigmas": (C.c_int, [_P, _F]),
    "zr_get_hotpix_info": (C.c_int, [_P, _D]),
    "zr_set_frc_translations": (C.c_int, [_P, _D, C.c_int]),
    "zr_get_frames": (C.c_int, [_P, _F]),
    "zr_pair_count": (C.c_int, [_P, _I]),
    "zr_equation_count": (C.c_int, [_P, _I]),
    "zr_get_solution": (C.c_int, [_P, _D, _D, _D, _D, _D, _D, _D]),
    "zr_get_sums": (C.c_int, [_P, _F, _F, _F, _F]),
    "zr_get_frc": (C.c_int, [_P, _D, _I]),
    "zr_device_results": (C.c_int, [_P, C.POINTER(_F), C.POINTER(_F), C.POINTER(_D)]),
    "zr_get_solver_info": (C.c_int, [_P, _I]),
    "zr_set_timing": (C.c_int, [_P, C.c_int]),
    "zr_get_stage_times": (C.c_int, [_P, _D, C.POINTER(C.c_long), C.c_int]),
    "zr_movie_info": (C.c_int, [_P, _I]),
    "zr_movie_device_bytes": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "zr_copy_sums_device": (C.c_int, [_P, _P, _P, _P, _P]),
    "zr_server_create": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(_P)]),
    "zr_server_create_ops": (C.c_int, [C.POINTER(ZrPlanOps), C.c_size_t, C.c_int, C.POINTER(_P)]),
    "zr_server_destroy": (None, [_P]),
    "zr_server_submit": (C.c_int, [_P, C.POINTER(ZrParams), C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_double,
                                   C.POINTER(C.c_longlong)]),
    "zr_server_warm": (C.c_int, [_P, C.POINTER(ZrParams), C.c_int, C.c_int, C.c_int, C.c_int, _I]),
    "zr_server_cancel": (C.c_int, [_P, C.c_longlong]),
    "zr_server_poll": (C.c_int, [_P, C.c_longlong, _I]),
    "zr_server_wait": (C.c_int, [_P, C.c_longlong, C.c_double, _I]),
    "zr_server_wait_any": (C.c_int, [_P, C.c_double, C.POINTER(C.c_longlong), _I]),
    "zr_server_result": (C.c_int, [_P, C.c_longlong, _D, C.c_int, _D, C.c_int, _I, _D, C.c_char_p, C.c_int]),
    "zr_server_release": (C.c_int, [_P, C.c_longlong]),
    "zr_server_stats": (C.c_int, [_P, C.POINTER(C.c_longlong)]),
    "zr_server_estimate_bytes": (C.c_size_t, [C.POINTER(ZrParams), C.c_int, C.c_int, C.c_int]),
}

STAGES = ["spectra", "pairs", "peaks", "solve", "shifts", "dose", "frc", "input", "tail"]

_lib = None


def header_symbols():
    """Function names declared in include/zorro_hip.h."""
    import re
    hdr = os.path.join(HERE, "..", "include", "zorro_hip.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(zr_[a-z_]+)\s*\(", txt)))


def load(path: str = LIB_PATH):
    """Load and type the library once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m zorro_amd.build` "
                          "(there is no CPU fallback for the HIP path)")
    # torch ships its own HIP runtime (same soname libamdhip64.so.7).  Loading
    # torch first makes the library bind to that one, so a process that moves
    # tensors with torch and runs these kernels has a single HIP runtime; the
    # other order loads two runtimes and torch then finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    if rc != ZR_OK:
        raise ZrError(rc, _lib.zr_last_error().decode())
    return rc


def fptr(a):
    """float* of a C-contiguous float32 numpy array."""
    return a.ctypes.data_as(_F)


def dptr(a):
    return a.ctypes.data_as(_D)
