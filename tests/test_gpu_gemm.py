"""GPU checks of the MFMA GEMM engine against numpy (float64 reference products).

FP64 tolerance: |C - C_ref| <= 1e-13 * (|A| |op(B)| + |beta C|) elementwise-bound scale; FP32:
2e-5 relative to the same scale (K up to 700 sums of products of O(1) numbers).
"""
import numpy as np
import pytest

from moao_amd.device import DeviceArray
from moao_amd.reconstructor import dgemm_dev, sgemm_dev

pytestmark = pytest.mark.gpu

CASES = [(128, 128, 16, 1, 0.0), (130, 77, 33, 1, 1.0), (300, 129, 200, 0, -1.0),
         (1, 1, 1, 0, 0.5), (257, 383, 700, 1, 1.0), (64, 500, 5, 0, 0.0)]


@pytest.mark.parametrize("M,N,K,transb,beta", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gemm_matches_numpy(M, N, K, transb, beta, dtype):
    print(f"[gemm] {dtype.__name__} M={M} N={N} K={K} transb={transb} beta={beta}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    lda, ldb, ldc = M + 3, (N if transb else K) + 5, M + 1
    A = np.asfortranarray(rng.standard_normal((lda, K)).astype(dtype))
    B = np.asfortranarray(rng.standard_normal((ldb, K if transb else N)).astype(dtype))
    C = np.asfortranarray(rng.standard_normal((ldc, N)).astype(dtype))
    alpha = -0.75
    dA, dB, dC = DeviceArray.from_host(A), DeviceArray.from_host(B), DeviceArray.from_host(C)
    f = dgemm_dev if dtype == np.float64 else sgemm_dev
    f(transb, M, N, K, alpha, dA.ptr, lda, dB.ptr, ldb, beta, dC.ptr, ldc)
    out = dC.to_host()
    a = A[:M].astype(np.float64)
    b = (B[:N].T if transb else B[:K]).astype(np.float64)
    ref = alpha * a @ b + beta * C[:M].astype(np.float64)
    scale = np.abs(alpha) * np.abs(a) @ np.abs(b) + abs(beta) * np.abs(C[:M])
    tol = 1e-13 if dtype == np.float64 else 2e-5
    assert np.all(np.abs(out[:M] - ref) <= tol * (scale + 1e-30))
    # rows beyond M (padding of ldc) untouched
    assert np.array_equal(out[M:], C[M:])
