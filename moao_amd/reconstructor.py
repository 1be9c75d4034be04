"""Reconstructor through the C-ABI: MOAO_LAPACK::reconstructor
(lapack/compute/reconstructor.cpp:32-58)."""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import check, lib


def reconstructor(nmeas: int, nmeasts: int, Cmm: np.ndarray, ldCmm: int, Ctm: np.ndarray,
                  ldCtm: int, ngal: int = 1) -> int:
    """Host-buffer drop-in, in place: Cmm -> L (lower), Ctm blocks -> R.  Returns the dpotrf
    info (0 on success).  Raises MoaoError on an argument/runtime error."""
    for a in (Cmm, Ctm):
        if a.dtype != np.float64 or not (a.flags.f_contiguous or a.ndim == 1):
            raise TypeError("Cmm and Ctm must be column-major float64")
    dp = ctypes.POINTER(ctypes.c_double)
    return check(lib().moao_hip_reconstructor(nmeas, nmeasts, Cmm.ctypes.data_as(dp), ldCmm,
                                              Ctm.ctypes.data_as(dp), ldCtm, ngal),
                 "moao_hip_reconstructor")


def sreconstructor(nmeas: int, nmeasts: int, Cmm: np.ndarray, ldCmm: int, Ctm: np.ndarray,
                   ldCtm: int, ngal: int = 1) -> int:
    """FP32 host drop-in (MOAO_CHAMELEON::reconstructor<float> arithmetic), in place."""
    for a in (Cmm, Ctm):
        if a.dtype != np.float32 or not (a.flags.f_contiguous or a.ndim == 1):
            raise TypeError("Cmm and Ctm must be column-major float32")
    fp = ctypes.POINTER(ctypes.c_float)
    return check(lib().moao_hip_sreconstructor(nmeas, nmeasts, Cmm.ctypes.data_as(fp), ldCmm,
                                               Ctm.ctypes.data_as(fp), ldCtm, ngal),
                 "moao_hip_sreconstructor")


def reconstructor_dev(nmeas: int, nmeasts: int, dCmm: int, ldCmm: int, dCtm: int, ldCtm: int,
                      ngal: int = 1, stream: int = 0) -> int:
    return check(lib().moao_hip_reconstructor_dev(nmeas, nmeasts, ctypes.c_void_p(dCmm), ldCmm,
                                                  ctypes.c_void_p(dCtm), ldCtm, ngal,
                                                  ctypes.c_void_p(stream)),
                 "moao_hip_reconstructor_dev")


def reconstructor_async(nmeas: int, nmeasts: int, dCmm: int, ldCmm: int, dCtm: int, ldCtm: int,
                        ngal: int, d_info: int, stream: int = 0) -> int:
    """Stream-ordered FP64 solve: returns once queued; the dpotrf status lands in the device int
    at d_info when the solve runs."""
    return check(lib().moao_hip_reconstructor_async(nmeas, nmeasts, ctypes.c_void_p(dCmm), ldCmm,
                                                    ctypes.c_void_p(dCtm), ldCtm, ngal,
                                                    ctypes.c_void_p(d_info),
                                                    ctypes.c_void_p(stream)),
                 "moao_hip_reconstructor_async")


def reconstructor_mixed_dev(nmeas: int, nmeasts: int, dCmm: int, ldCmm: int, dCtm: int,
                            ldCtm: int, tomo=None, tol: float = 1e-11, max_iter: int = 5,
                            stream: int = 0, allow_noconv: bool = False, mode=None):
    """Mixed-precision solve (FP32 or split-BF16 factorization + FP64 refinement), Ctm -> R in
    place, Cmm preserved.  Returns (status, relative residual, refinement steps)."""
    from .lib import MOAO_ERR_NOCONV, MOAO_MIXED_FP32
    mode = MOAO_MIXED_FP32 if mode is None else mode
    view = tomo.view() if tomo is not None else None
    res = ctypes.c_double()
    its = ctypes.c_int()
    rc = lib().moao_hip_reconstructor_mixed_dev(
        nmeas, nmeasts, ctypes.c_void_p(dCmm), ldCmm, ctypes.c_void_p(dCtm), ldCtm,
        ctypes.byref(view) if view is not None else None, mode, tol, max_iter,
        ctypes.byref(res), ctypes.byref(its), ctypes.c_void_p(stream))
    if not (allow_noconv and rc == MOAO_ERR_NOCONV):
        check(rc, "moao_hip_reconstructor_mixed_dev")
    return rc, res.value, its.value


def dgemm_dev(transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, stream=0):
    check(lib().moao_hip_dgemm_dev(transb, M, N, K, alpha, ctypes.c_void_p(dA), lda,
                                   ctypes.c_void_p(dB), ldb, beta, ctypes.c_void_p(dC), ldc,
                                   ctypes.c_void_p(stream)), "moao_hip_dgemm_dev")


def sgemm_dev(transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, stream=0):
    check(lib().moao_hip_sgemm_dev(transb, M, N, K, alpha, ctypes.c_void_p(dA), lda,
                                   ctypes.c_void_p(dB), ldb, beta, ctypes.c_void_p(dC), ldc,
                                   ctypes.c_void_p(stream)), "moao_hip_sgemm_dev")


def sgemm_bf16x3_dev(transb, M, N, K, alpha, dA, lda, dB, ldb, beta, dC, ldc, stream=0):
    check(lib().moao_hip_sgemm_bf16x3_dev(transb, M, N, K, alpha, ctypes.c_void_p(dA), lda,
                                          ctypes.c_void_p(dB), ldb, beta, ctypes.c_void_p(dC),
                                          ldc, ctypes.c_void_p(stream)),
          "moao_hip_sgemm_bf16x3_dev")
