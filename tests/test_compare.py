"""The library's compareMatrices (moao_compare_matrices on the host, moao_hip_compare_matrices_dev
on device buffers) against the oracle's restatement (oracle/check.py, itself pinned to the
hand-built reference's compareMatrices): every uplo / trans rule, zeros in A (the |B| branch),
non-square shapes.  max_err is order-free and must match exactly; the Frobenius sums of squares
are accumulated in different orders (sequential on the host, per-workgroup partials on the
device), so they agree within the n u bound of a sum of n positive terms: 1e-13 relative for
the small cases, 1e-9 for the 7.5M-entry one."""
import ctypes

import numpy as np
import pytest

from moao_amd.lib import lib
from oracle.check import compare_matrices

dp = ctypes.POINTER(ctypes.c_double)
CASES = [(37, 37, u, t) for u in "NUL" for t in "NT"] + [(23, 41, "N", "N"), (41, 23, "N", "T"),
                                                          (30, 30, "X", "T")]


def _inputs(nr, nc, seed):
    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.standard_normal((nr, nc)))
    A[rng.random((nr, nc)) < 0.1] = 0.0
    B = A.reshape(-1, order="F") * (1 + 1e-6 * rng.standard_normal(nr * nc))
    B[rng.random(nr * nc) < 0.05] += 1.0
    return A, np.ascontiguousarray(B)


@pytest.mark.parametrize("nr,nc,uplo,trans", CASES)
def test_host_compare_matches_oracle(nr, nc, uplo, trans):
    A, B = _inputs(nr, nc, nr * 7 + nc)
    mx, fr = ctypes.c_double(), ctypes.c_double()
    assert lib().moao_compare_matrices(A.ctypes.data_as(dp), B.ctypes.data_as(dp), nr, nc,
                                       uplo.encode(), trans.encode(), ctypes.byref(mx),
                                       ctypes.byref(fr)) == 0
    omx, ofr = compare_matrices(A, B, nr, nc, uplo, trans)
    assert mx.value == omx
    assert abs(fr.value - ofr) <= 1e-13 * ofr


@pytest.mark.gpu
@pytest.mark.parametrize("nr,nc,uplo,trans", CASES + [(3000, 2500, "N", "T"), (2000, 2000, "L", "T")])
def test_device_compare_matches_host(nr, nc, uplo, trans):
    from moao_amd.device import DeviceArray
    A, B = _inputs(nr, nc, nr + nc)
    dA, dB = DeviceArray.from_host(A), DeviceArray.from_host(B)
    mx, fr, hmx, hfr = (ctypes.c_double() for _ in range(4))
    assert lib().moao_hip_compare_matrices_dev(ctypes.c_void_p(dA.ptr), ctypes.c_void_p(dB.ptr),
                                               nr, nc, uplo.encode(), trans.encode(),
                                               ctypes.byref(mx), ctypes.byref(fr), None) == 0
    lib().moao_compare_matrices(A.ctypes.data_as(dp), B.ctypes.data_as(dp), nr, nc,
                                uplo.encode(), trans.encode(), ctypes.byref(hmx), ctypes.byref(hfr))
    assert mx.value == hmx.value
    assert abs(fr.value - hfr.value) <= max(1e-13, 2e-16 * nr * nc) * hfr.value
