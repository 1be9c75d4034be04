"""matcov through the C-ABI: MOAO_COMMON::matcov_tile (common/compute/matcov.cpp:17-98)."""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import check, lib
from .tomo import Tomo

CMM, CMAA, CTM, CTT = 1, 2, 3, 4


def matcov_tile(data: np.ndarray, nrows: int, ncols: int, xoffset: int, yoffset: int, lda: int,
                tomo: Tomo, type_mat: int) -> None:
    """Host-buffer drop-in: generates the tile on the GPU and writes it into ``data``
    (float64, column-major with leading dimension lda), in place."""
    if data.dtype != np.float64:
        raise TypeError("data must be float64")
    if not (data.flags.f_contiguous or data.ndim == 1):
        raise TypeError("data must be column-major")
    if ncols and data.size < (ncols - 1) * lda + nrows:
        raise ValueError("data too small for (nrows, ncols, lda)")
    v = tomo.view()
    check(lib().moao_hip_matcov_tile(data.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     nrows, ncols, xoffset, yoffset, lda, ctypes.byref(v),
                                     type_mat), "moao_hip_matcov_tile")


def matcov_tile_dev(d_ptr: int, nrows: int, ncols: int, xoffset: int, yoffset: int, lda: int,
                    tomo: Tomo, type_mat: int, stream: int = 0) -> None:
    v = tomo.view()
    check(lib().moao_hip_matcov_tile_dev(ctypes.c_void_p(d_ptr), nrows, ncols, xoffset, yoffset,
                                         lda, ctypes.byref(v), type_mat,
                                         ctypes.c_void_p(stream)), "moao_hip_matcov_tile_dev")


def cmm(tomo: Tomo) -> np.ndarray:
    n = tomo.nmeas()
    a = np.zeros((n, n), order="F")
    matcov_tile(a, n, n, 0, 0, n, tomo, CMM)
    return a


def ctm(tomo: Tomo) -> np.ndarray:
    n, m = tomo.nmeas(), tomo.nmeasts()
    a = np.zeros((m, n), order="F")
    matcov_tile(a, m, n, 0, 0, m, tomo, CTM)
    return a


def ctt(tomo: Tomo) -> np.ndarray:
    m = tomo.nmeasts()
    a = np.zeros((m, m), order="F")
    matcov_tile(a, m, m, 0, 0, m, tomo, CTT)
    return a
