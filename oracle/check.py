"""Error metrics (oracle side; test infrastructure only).

compare_matrices restates compareMatrices (common/compute/check_utils.cpp:20-57):
  maxErr  = max over the visited entries of |1 - B/A| (|B| where A == 0);
  frobN   = sqrt(sum (B - A)^2), off-diagonal terms counted twice when uplo is 'U' or 'L'.
A is column-major nrows x ncols; B is read transposed when trans == 'T' (a row-major file).
With uplo 'L' the reference visits j < i (strict lower), with 'U' j >= i.
Pinned by tests/test_oracle_tomo.py against the hand-built reference's compareMatrices.
"""
from __future__ import annotations

import numpy as np


def compare_matrices(A: np.ndarray, B: np.ndarray, nrows: int, ncols: int, uplo: str = "N",
                     trans: str = "N"):
    a = np.asarray(A, dtype=np.float64).reshape(-1, order="F")
    b = np.asarray(B, dtype=np.float64).reshape(-1, order="F")
    if uplo not in ("U", "L"):
        uplo = "N"
    i = np.arange(nrows)[:, None]
    j = np.arange(ncols)[None, :]
    if uplo == "U":
        mask = j >= i
    elif uplo == "L":
        mask = j < i
    else:
        mask = np.ones((nrows, ncols), dtype=bool)
    ii, jj = np.nonzero(mask)
    nA = ii + jj * nrows
    nB = nrows * jj + ii if trans == "T" else ncols * ii + jj
    av, bv = a[nA], b[nB]
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(av == 0.0, -bv, 1.0 - bv / np.where(av == 0.0, 1.0, av))
    max_err = float(np.max(np.abs(d))) if d.size else 0.0
    w = np.where((uplo == "N") | (ii == jj), 1.0, 2.0)
    frob = float(np.sqrt(np.sum((bv - av) ** 2 * w)))
    return max_err, frob


def rel_fro(X: np.ndarray, Xref: np.ndarray) -> float:
    return float(np.linalg.norm(X - Xref) / np.linalg.norm(Xref))
