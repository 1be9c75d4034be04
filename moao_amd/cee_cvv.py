"""Cee/Cvv through the C-ABI: MOAO_LAPACK::Cee_Cvv (lapack/compute/cee_cvv.cpp:50-124)."""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import check, lib


def cee_cvv(nmeas, nmeasts, nact, Cmm, Cpp, Cpm, R, Dx, Cee, Cvv, Tmp):
    """Host-buffer drop-in with natural leading dimensions; Cpp/Cpm/Cee/Cvv/Tmp in place."""
    dp = ctypes.POINTER(ctypes.c_double)
    arrs = (Cmm, Cpp, Cpm, R, Dx, Cee, Cvv, Tmp)
    for a in arrs:
        if a.dtype != np.float64 or not a.flags.f_contiguous:
            raise TypeError("all operands must be column-major float64")
    lds = (nmeas, nmeasts, nmeasts, nmeasts, nact, nmeasts, nact, nact)
    args = []
    for a, ld in zip(arrs, lds):
        args += [a.ctypes.data_as(dp), max(1, ld)]
    check(lib().moao_hip_cee_cvv(nmeas, nmeasts, nact, *args), "moao_hip_cee_cvv")
