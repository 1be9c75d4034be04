"""ctypes front of the oracle's C matcov restatement (oracle/matcov_oracle.c).

Test infrastructure only.  See matcov_oracle.c for the reference lines each function follows.
Pinned by tests/test_oracle_matcov.py: bit-for-bit equality with golden tiles written by the
hand-built reference, the series constants against scipy.special.gamma, D(r) against the exact
von Karman structure function built on scipy.special.kv, its L0 -> infinity Kolmogorov limit
6.88 r^(5/3), and matrix invariants (symmetry, positive definiteness, translation invariance).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmoao_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


class OracleTomoC(ctypes.Structure):
    _fields_ = [("Nw", ctypes.c_int), ("Nlayer", ctypes.c_int), ("nlgs", ctypes.c_int),
                ("Nx", ctypes.c_long), ("diam", ctypes.c_double), ("lgs_cst", ctypes.c_double),
                ("Nsubap", _lp), ("Nssp", _lp),
                ("sspSizeL", _dp), ("u", _dp), ("v", _dp), ("X", _dp), ("Y", _dp),
                ("L0", _dp), ("cn2", _dp),
                ("noiseNGS", _dp), ("noiseLGSxx", _dp), ("noiseLGSyy", _dp),
                ("noiseLGSxy", _dp)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_matcov_tile.restype = ctypes.c_int
        _lib.oracle_matcov_tile.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(OracleTomoC), ctypes.c_int]
        for f in ("vk_dphi",):
            getattr(_lib, f).restype = ctypes.c_double
            getattr(_lib, f).argtypes = [ctypes.c_double] * 3
        for f in ("vk_series", "vk_asymptote"):
            getattr(_lib, f).restype = ctypes.c_double
            getattr(_lib, f).argtypes = [ctypes.c_double]
    return _lib


def dphi(x, y, L0):
    return lib().vk_dphi(float(x), float(y), float(L0))


def _view(t):
    keep = {
        "Nsubap": np.ascontiguousarray(t.Nsubap, dtype=np.int64),
        "Nssp": np.ascontiguousarray(t.Nssp, dtype=np.int64),
        "sspSizeL": np.ascontiguousarray(t.sspSizeL, dtype=np.float64),
        "u": np.ascontiguousarray(t.u, dtype=np.float64),
        "v": np.ascontiguousarray(t.v, dtype=np.float64),
        "X": np.ascontiguousarray(t.X, dtype=np.float64),
        "Y": np.ascontiguousarray(t.Y, dtype=np.float64),
        "L0": np.ascontiguousarray(t.L0diff, dtype=np.float64),
        "cn2": np.ascontiguousarray(t.cn2, dtype=np.float64),
        "noiseNGS": np.ascontiguousarray(t.noiseNGS, dtype=np.float64),
        "noiseLGSxx": np.ascontiguousarray(t.noiseLGSxx, dtype=np.float64),
        "noiseLGSyy": np.ascontiguousarray(t.noiseLGSyy, dtype=np.float64),
        "noiseLGSxy": np.ascontiguousarray(t.noiseLGSxy, dtype=np.float64),
    }
    c = OracleTomoC()
    c.Nw = t.sys.nW
    c.Nlayer = t.atm.nLayer
    c.nlgs = t.sys.nLgs
    c.Nx = t.Nx
    c.diam = t.sys.diam
    c.lgs_cst = t.sys.lgsCst
    for k, a in keep.items():
        setattr(c, k, a.ctypes.data_as(_lp if a.dtype == np.int64 else _dp))
    return c, keep


def matcov_tile(data: np.ndarray, nrows, ncols, xoffset, yoffset, lda, tomo, type_mat) -> int:
    """Fill the column-major buffer ``data`` (float64, F-ordered or flat) like
    MOAO_COMMON::matcov_tile (common/compute/matcov.cpp:17-98)."""
    assert data.dtype == np.float64
    c, keep = _view(tomo)
    return lib().oracle_matcov_tile(data.ctypes.data_as(_dp), nrows, ncols, xoffset, yoffset,
                                    lda, ctypes.byref(c), type_mat)


def cmm(tomo) -> np.ndarray:
    n = tomo.nmeas()
    a = np.zeros((n, n), order="F")
    matcov_tile(a, n, n, 0, 0, n, tomo, 1)
    return a


def ctm(tomo) -> np.ndarray:
    n, m = tomo.nmeas(), tomo.nmeasts()
    a = np.zeros((m, n), order="F")
    matcov_tile(a, m, n, 0, 0, m, tomo, 3)
    return a


def ctt(tomo) -> np.ndarray:
    m = tomo.nmeasts()
    a = np.zeros((m, m), order="F")
    matcov_tile(a, m, m, 0, 0, m, tomo, 4)
    return a
