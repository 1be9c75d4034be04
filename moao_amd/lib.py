"""Loader for libmoao_hip.so (the C-ABI declared in include/moao_hip.h).

The product path has no fallback: if the shared library is missing or fails to load, every
entry point raises.  Build it with ``python -m moao_amd.build`` (or __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmoao_hip.so")

MOAO_OK = 0
MOAO_ERR_ARG = -1
MOAO_ERR_HIP = -2
MOAO_ERR_NOMEM = -3
MOAO_ERR_NOCONV = -4
MOAO_ERR_IO = -5
MOAO_ERR_NODEV = -6

MOAO_FP64 = 0
MOAO_MIXED_FP32 = 1
MOAO_MIXED_BF16 = 2

_c_long_p = ctypes.POINTER(ctypes.c_long)
_c_double_p = ctypes.POINTER(ctypes.c_double)


class MoaoError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


class TomoView(ctypes.Structure):
    """ctypes mirror of moao_tomo_view (include/moao_hip.h)."""
    _fields_ = [("Nw", ctypes.c_int), ("Nlayer", ctypes.c_int), ("nlgs", ctypes.c_int),
                ("Nx", ctypes.c_long), ("diam", ctypes.c_double), ("lgs_cst", ctypes.c_double),
                ("Nsubap", _c_long_p), ("Nssp", _c_long_p),
                ("sspSizeL", _c_double_p), ("u", _c_double_p), ("v", _c_double_p),
                ("X", _c_double_p), ("Y", _c_double_p), ("L0", _c_double_p),
                ("cn2", _c_double_p), ("noiseNGS", _c_double_p), ("noiseLGSxx", _c_double_p),
                ("noiseLGSyy", _c_double_p), ("noiseLGSxy", _c_double_p)]


class ServeConfig(ctypes.Structure):
    """ctypes mirror of moao_serve_config."""
    _fields_ = [("slots", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("max_bypass", ctypes.c_int)]


class ServeRequest(ctypes.Structure):
    """ctypes mirror of moao_serve_request."""
    _fields_ = [("sys_file", ctypes.c_char_p), ("atm_file", ctypes.c_char_p),
                ("target", ctypes.c_int), ("mode", ctypes.c_int), ("tol", ctypes.c_double),
                ("max_iter", ctypes.c_int), ("R", _c_double_p), ("ldR", ctypes.c_long),
                ("R_cols", ctypes.c_long)]


class ServeResult(ctypes.Structure):
    """ctypes mirror of moao_serve_result."""
    _fields_ = [("status", ctypes.c_int), ("slot", ctypes.c_int), ("n", ctypes.c_long),
                ("m", ctypes.c_long), ("iters", ctypes.c_int), ("resid", ctypes.c_double),
                ("footprint", ctypes.c_size_t), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_end", ctypes.c_double)]


class IsampleStruct(ctypes.Structure):
    """ctypes mirror of moao_isample_struct (the reference's isample_struct fields)."""
    _fields_ = [("N", ctypes.c_long), ("np", ctypes.c_long), ("nidx", ctypes.c_long),
                ("lambda_", ctypes.c_double), ("dactupix", ctypes.c_double),
                ("idx", _c_long_p), ("abs2fi", _c_double_p), ("otf_tel", _c_double_p),
                ("dphi", _c_double_p), ("filename_idx", ctypes.c_char * 256),
                ("filename_abs2fi", ctypes.c_char * 256),
                ("filename_otftel", ctypes.c_char * 256), ("precision", ctypes.c_int),
                ("dev", ctypes.c_void_p)]


class TileDesc(ctypes.Structure):
    """ctypes mirror of moao_hip_desc."""
    _fields_ = [("dtype", ctypes.c_int), ("mat", ctypes.c_void_p), ("mb", ctypes.c_int),
                ("nb", ctypes.c_int), ("lm", ctypes.c_long), ("ln", ctypes.c_long)]


class ServeCounters(ctypes.Structure):
    """ctypes mirror of moao_serve_counters."""
    _fields_ = [("queued", ctypes.c_long), ("running", ctypes.c_long),
                ("completed", ctypes.c_long), ("peak_running", ctypes.c_long),
                ("bytes_in_use", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t),
                ("budget", ctypes.c_size_t), ("graph_captures", ctypes.c_long),
                ("graph_replays", ctypes.c_long)]


_lib = None

# name -> (restype, argtypes); the exported C-ABI
_SIGS = {
    "moao_last_error": (ctypes.c_char_p, []),
    "moao_version": (ctypes.c_char_p, []),
    "moao_tomo_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "moao_tomo_update_atm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "moao_tomo_set_target": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "moao_tomo_set_target_coord": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int]),
    "moao_tomo_get_view": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TomoView)]),
    "moao_tomo_nmeas": (ctypes.c_long, [ctypes.c_void_p]),
    "moao_tomo_nmeasts": (ctypes.c_long, [ctypes.c_void_p]),
    "moao_tomo_dims": (ctypes.c_int, [ctypes.c_void_p, _c_long_p]),
    "moao_tomo_field": (ctypes.c_long, [ctypes.c_void_p, ctypes.c_char_p, _c_double_p,
                                        ctypes.c_long]),
    "moao_tomo_destroy": (None, [ctypes.c_void_p]),
    "moao_hip_init": (ctypes.c_int, [ctypes.c_int]),
    "moao_hip_finalize": (ctypes.c_int, []),
    "moao_hip_malloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    "moao_hip_free": (ctypes.c_int, [ctypes.c_void_p]),
    "moao_hip_memcpy_h2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "moao_hip_memcpy_d2h": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "moao_hip_memcpy_d2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "moao_hip_synchronize": (ctypes.c_int, []),
    "moao_hip_mem_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.POINTER(ctypes.c_size_t)]),
    "moao_hip_stream_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "moao_hip_stream_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "moao_hip_stream_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "moao_hip_matcov_tile": (ctypes.c_int, [_c_double_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(TomoView), ctypes.c_int]),
    "moao_hip_matcov_tile_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.POINTER(TomoView), ctypes.c_int,
                                                ctypes.c_void_p]),
    "moao_hip_reconstructor": (ctypes.c_int, [ctypes.c_long, ctypes.c_long, _c_double_p,
                                              ctypes.c_long, _c_double_p, ctypes.c_long,
                                              ctypes.c_long]),
    "moao_hip_reconstructor_dev": (ctypes.c_int, [ctypes.c_long, ctypes.c_long, ctypes.c_void_p,
                                                  ctypes.c_long, ctypes.c_void_p, ctypes.c_long,
                                                  ctypes.c_long, ctypes.c_void_p]),
    "moao_hip_sreconstructor": (ctypes.c_int, [ctypes.c_long, ctypes.c_long,
                                               ctypes.POINTER(ctypes.c_float), ctypes.c_long,
                                               ctypes.POINTER(ctypes.c_float), ctypes.c_long,
                                               ctypes.c_long]),
    "moao_hip_sreconstructor_dev": (ctypes.c_int, [ctypes.c_long, ctypes.c_long,
                                                   ctypes.c_void_p, ctypes.c_long,
                                                   ctypes.c_void_p, ctypes.c_long, ctypes.c_long,
                                                   ctypes.c_void_p]),
    "moao_hip_reconstructor_async": (ctypes.c_int, [ctypes.c_long, ctypes.c_long,
                                                    ctypes.c_void_p, ctypes.c_long,
                                                    ctypes.c_void_p, ctypes.c_long,
                                                    ctypes.c_long, ctypes.c_void_p,
                                                    ctypes.c_void_p]),
    "moao_hip_reconstructor_noalloc": (ctypes.c_int, [ctypes.c_long, ctypes.c_long,
                                                      ctypes.c_void_p, ctypes.c_long,
                                                      ctypes.c_void_p, ctypes.c_long,
                                                      ctypes.c_long, ctypes.c_void_p,
                                                      ctypes.c_void_p]),
    "moao_hip_solve_graph_create": (ctypes.c_int, [ctypes.c_long, ctypes.c_long, ctypes.c_void_p,
                                                   ctypes.c_long, ctypes.c_void_p, ctypes.c_long,
                                                   ctypes.c_long, ctypes.c_void_p,
                                                   ctypes.POINTER(ctypes.c_void_p)]),
    "moao_hip_solve_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "moao_hip_solve_graph_destroy": (None, [ctypes.c_void_p]),
    "moao_hip_sequence_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "moao_hip_sequence_wait": (ctypes.c_int, [ctypes.c_void_p]),
    "moao_hip_sequence_destroy": (None, [ctypes.c_void_p]),
    "moao_hip_lapack_to_tile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_long,
                                               ctypes.POINTER(TileDesc), ctypes.c_void_p]),
    "moao_hip_tile_to_lapack": (ctypes.c_int, [ctypes.POINTER(TileDesc), ctypes.c_void_p,
                                               ctypes.c_long, ctypes.c_void_p]),
    "moao_hip_reconstructor_tile_async": (ctypes.c_int, [ctypes.POINTER(TileDesc),
                                                         ctypes.POINTER(TileDesc),
                                                         ctypes.c_void_p, ctypes.c_int,
                                                         ctypes.c_int]),
    "moao_hip_reconstructor_mixed_dev": (ctypes.c_int, [ctypes.c_long, ctypes.c_long,
                                                        ctypes.c_void_p, ctypes.c_long,
                                                        ctypes.c_void_p, ctypes.c_long,
                                                        ctypes.POINTER(TomoView), ctypes.c_int,
                                                        ctypes.c_double, ctypes.c_int,
                                                        _c_double_p,
                                                        ctypes.POINTER(ctypes.c_int),
                                                        ctypes.c_void_p]),
    "moao_hip_dgemm_dev": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_void_p, ctypes.c_long,
                                          ctypes.c_void_p, ctypes.c_long, ctypes.c_double,
                                          ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p]),
    "moao_hip_sgemm_dev": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_void_p, ctypes.c_long,
                                          ctypes.c_void_p, ctypes.c_long, ctypes.c_float,
                                          ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p]),
    "moao_hip_sgemm_bf16x3_dev": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                                 ctypes.c_long, ctypes.c_void_p, ctypes.c_long,
                                                 ctypes.c_float, ctypes.c_void_p, ctypes.c_long,
                                                 ctypes.c_void_p]),
    "moao_hip_cee_cvv": (ctypes.c_int, [ctypes.c_long, ctypes.c_long, ctypes.c_long]
                         + [_c_double_p, ctypes.c_long] * 8),
    "moao_hip_cee_cvv_dev": (ctypes.c_int, [ctypes.c_long, ctypes.c_long, ctypes.c_long]
                             + [ctypes.c_void_p, ctypes.c_long] * 8 + [ctypes.c_void_p]),
    "moao_hip_cee_cvv_batched_dev": (ctypes.c_int, [ctypes.c_long] * 4
                                     + [ctypes.c_void_p, ctypes.c_long] * 8 + [ctypes.c_void_p]),
    "moao_intersample_default_n": (ctypes.c_long, [ctypes.c_long]),
    "moao_hip_intersample_create": (ctypes.c_int, [ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                                   _c_long_p, _c_double_p, _c_double_p,
                                                   ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_void_p)]),
    "moao_hip_intersample_n": (ctypes.c_long, [ctypes.c_void_p]),
    "moao_hip_intersample_process": (ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    "moao_hip_intersample_process_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_int, ctypes.c_void_p,
                                                        ctypes.c_void_p]),
    "moao_hip_intersample_destroy": (None, [ctypes.c_void_p]),
    "moao_hip_psf_accumulate_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_long,
                                                   ctypes.c_void_p, ctypes.c_void_p]),
    "moao_hip_intersample_prepare": (None, [ctypes.POINTER(IsampleStruct), ctypes.c_long,
                                            ctypes.c_long, ctypes.c_float, ctypes.c_char_p]),
    "moao_hip_intersample_init": (ctypes.c_int, [ctypes.POINTER(IsampleStruct)]),
    "moao_hip_intersample_process_d": (ctypes.c_int, [ctypes.POINTER(IsampleStruct),
                                                      _c_double_p]),
    "moao_hip_intersample_process_f": (ctypes.c_int, [ctypes.POINTER(IsampleStruct),
                                                      ctypes.POINTER(ctypes.c_float)]),
    "moao_hip_intersample_free": (None, [ctypes.POINTER(IsampleStruct)]),
    "moao_compare_matrices": (ctypes.c_int, [_c_double_p, _c_double_p, ctypes.c_long,
                                             ctypes.c_long, ctypes.c_char, ctypes.c_char,
                                             _c_double_p, _c_double_p]),
    "moao_hip_compare_matrices_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_long, ctypes.c_long, ctypes.c_char,
                                                     ctypes.c_char, _c_double_p, _c_double_p,
                                                     ctypes.c_void_p]),
    "moao_fits_read": (ctypes.c_int, [ctypes.c_char_p, _c_long_p, _c_long_p, _c_double_p,
                                      ctypes.c_long]),
    "moao_fits_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_long, ctypes.c_long,
                                       _c_double_p]),
    "moao_hip_profile_begin": (ctypes.c_int, []),
    "moao_hip_profile_end": (ctypes.c_int, [_c_double_p, _c_double_p, _c_long_p]),
    "moao_serve_create": (ctypes.c_int, [ctypes.POINTER(ServeConfig),
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "moao_serve_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServeRequest),
                                         _c_long_p]),
    "moao_serve_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_long,
                                       ctypes.POINTER(ServeResult)]),
    "moao_serve_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServeCounters)]),
    "moao_serve_destroy": (None, [ctypes.c_void_p]),
}

PROF_CLASSES = ["gemm_f64", "gemm_f32", "leaf", "matcov", "convert", "fft", "other"]


def lib() -> ctypes.CDLL:
    """Load libmoao_hip.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m moao_amd.build` (no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(rc: int, where: str) -> int:
    if rc < 0:
        msg = lib().moao_last_error()
        raise MoaoError(rc, where, msg.decode() if msg else "")
    return rc
