"""The reference's lapack path restated with host LAPACK/BLAS (oracle; test infrastructure).

reconstructor  lapack/compute/reconstructor.cpp:32-58
    dpotrf('L', Cmm); for each target g: Ctm_g = &Ctm[n*m*g];
    dtrsm(Right, Lower, Trans, NonUnit, 1.0, Cmm, Ctm_g); dtrsm(Right, Lower, NoTrans, ...).
Cee_Cvv        lapack/compute/cee_cvv.cpp:50-124 (+ geadd 144-161)
    Cpp -= Cpm R^T + R Cpm^T (dsyr2k, lower);  Cpm = R Cmm (dsymm right lower);
    Cee = lower(Cpp) mirrored (dlaset U + dlacpy L + in-place geadd('T') + dscal(0.5) on the
    diagonal);  Cee += Cpm R^T (dgemm N,T);  Tmp = Dx Cee (dsymm right lower);
    Cvv = Tmp Dx^T (dgemm N,T).
The reference passes nmeasts/nmeas/nact as leading dimensions regardless of its ld* arguments;
these functions take plain column-major numpy arrays (leading dimension = rows).

The LAPACK/BLAS underneath is scipy's (OpenBLAS); the reference's CI used OpenBLAS 0.2.20 or
MKL 2018.  Pinned by: tests/test_oracle_lapack.py -- residual ||Ctm - R Cmm|| / ||Ctm|| at
FP64 round-off, agreement with scipy.linalg.cho_solve (a textbook routine), the closed form
Cee = Ctt - Ctm R^T - R Ctm^T + R Cmm R^T, symmetry of Cee/Cvv, and the geadd mirror rule on a
hand-checked 3x3 case.  The same functions, timed, are bench.py's CPU baseline.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import blas, lapack


def reconstructor(Cmm: np.ndarray, Ctm: np.ndarray, ngal: int = 1):
    """In place like the reference: Cmm (n x n, F-order) -> L in its lower triangle; Ctm
    (m x n*ngal... as ngal consecutive m x n blocks in a flat F-order buffer, or one m x n
    array) -> R.  Returns LAPACK info (0 = success; R untouched otherwise)."""
    n = Cmm.shape[0]
    c, info = lapack.dpotrf(Cmm, lower=1, clean=0, overwrite_a=1)
    if c is not Cmm:
        Cmm[...] = c
    if info != 0:
        return int(info)
    blocks = _targets(Ctm, n, ngal)
    for b in blocks:
        x = blas.dtrsm(1.0, Cmm, b, side=1, lower=1, trans_a=1, diag=0, overwrite_b=1)
        x = blas.dtrsm(1.0, Cmm, x, side=1, lower=1, trans_a=0, diag=0, overwrite_b=1)
        if x is not b:
            b[...] = x
    return 0


def _targets(Ctm, n, ngal):
    if ngal == 1 and Ctm.ndim == 2:
        return [Ctm]
    flat = Ctm.reshape(-1, order="F")
    m = flat.size // (n * ngal)
    return [flat[g * n * m:(g + 1) * n * m].reshape((m, n), order="F") for g in range(ngal)]


def geadd_T_inplace(B: np.ndarray):
    """MOAO_LAPACK::geadd('T', M, N, 1, B, ld, 1, B, ld) with A aliasing B
    (lapack/compute/cee_cvv.cpp:155-159): i outer, j inner, B(j,i) <- B(i,j) + B(j,i)."""
    M, N = B.shape
    flat = B.reshape(-1, order="F")
    ld = M
    for i in range(M):
        for j in range(N):
            flat[i * ld + j] = flat[i + j * ld] + flat[i * ld + j]


def cee_cvv(Cmm, Cpp, Cpm, R, Dx, small_geadd: bool = False):
    """Returns (Cee, Cvv) and updates Cpp / Cpm in place as the reference does."""
    m = Cpp.shape[0]
    r = blas.dsyr2k(-1.0, Cpm, R, beta=1.0, c=Cpp, trans=0, lower=1, overwrite_c=1)
    if r is not Cpp:
        Cpp[...] = r
    r = blas.dsymm(1.0, Cmm, R, beta=0.0, c=Cpm, side=1, lower=1, overwrite_c=1)
    if r is not Cpm:
        Cpm[...] = r
    Cee = np.zeros((m, m), order="F")                   # laset('U', 0, 0)
    Cee[np.tril_indices(m)] = Cpp[np.tril_indices(m)]   # lacpy('L')
    if small_geadd:
        geadd_T_inplace(Cee)                            # literal loop (slow; small m only)
    else:
        # the loop's effect: strictly-upper <- strictly-lower mirrored, diagonal doubled
        low = np.tril(Cee, -1)
        Cee = np.asfortranarray(low + low.T + 2.0 * np.diag(np.diag(Cee)))
    Cee[np.diag_indices(m)] *= 0.5                      # dscal(0.5) on the diagonal
    Cee = blas.dgemm(1.0, Cpm, R, beta=1.0, c=Cee, trans_a=0, trans_b=1, overwrite_c=1)
    Tmp = blas.dsymm(1.0, Cee, Dx, beta=0.0, side=1, lower=1)
    Cvv = blas.dgemm(1.0, Tmp, Dx, beta=0.0, trans_a=0, trans_b=1)
    return np.asfortranarray(Cee), np.asfortranarray(Cvv)
