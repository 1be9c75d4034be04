"""GPU checks of the MFMA GEMM engine against numpy (float64 reference products).

FP64 tolerance: |C - C_ref| <= 1e-13 * (|A| |op(B)| + |beta C|) elementwise-bound scale; FP32:
2e-5 relative to the same scale, widened in proportion to K beyond 700 (sums of K products of
O(1) numbers).  Shapes cover both tile sizes (launches of >= 512 128-row tiles and smaller
ones), ragged M, N and K (K not a multiple of the 8-deep slice), and transb = 0, 1 and 2
(symmetric B read from its lower triangle).
"""
import numpy as np
import pytest

from moao_amd.device import DeviceArray
from moao_amd.reconstructor import dgemm_dev, sgemm_dev

pytestmark = pytest.mark.gpu

CASES = [(128, 128, 16, 1, 0.0), (130, 77, 33, 1, 1.0), (300, 129, 200, 0, -1.0),
         (1, 1, 1, 0, 0.5), (257, 383, 700, 1, 1.0), (64, 500, 5, 0, 0.0),
         # >= 512 tiles of 128 rows: the 128-row kernels; 561 tiles leave a short last wave
         # whose tiles run as K chunks plus a reduction
         (4133, 3900, 300, 1, 1.0), (4133, 3900, 301, 0, -1.0), (4133, 2100, 301, 0, 1.0),
         # symmetric B (K == N)
         (50, 300, 300, 2, 1.0), (257, 129, 129, 2, 0.0), (7, 16, 16, 2, 1.0),
         (200, 77, 77, 2, -1.0), (4133, 2100, 2100, 2, 0.5)]


def sym_lower(B, K):
    L = np.tril(B[:K, :K].astype(np.float64))
    return L + np.tril(L, -1).T


def op_b(B, transb, N, K):
    if transb == 1:
        return B[:N].T.astype(np.float64)
    if transb == 0:
        return B[:K].astype(np.float64)
    return sym_lower(B, K)


@pytest.mark.parametrize("M,N,K,transb,beta", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gemm_matches_numpy(M, N, K, transb, beta, dtype):
    print(f"[gemm] {dtype.__name__} M={M} N={N} K={K} transb={transb} beta={beta}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    lda, ldb, ldc = M + 3, (N if transb == 1 else K) + 5, M + 1
    A = np.asfortranarray(rng.standard_normal((lda, K)).astype(dtype))
    B = np.asfortranarray(rng.standard_normal((ldb, N if transb == 0 else K)).astype(dtype))
    C = np.asfortranarray(rng.standard_normal((ldc, N)).astype(dtype))
    alpha = -0.75
    dA, dB, dC = DeviceArray.from_host(A), DeviceArray.from_host(B), DeviceArray.from_host(C)
    f = dgemm_dev if dtype == np.float64 else sgemm_dev
    f(transb, M, N, K, alpha, dA.ptr, lda, dB.ptr, ldb, beta, dC.ptr, ldc)
    out = dC.to_host()
    a = A[:M].astype(np.float64)
    b = op_b(B, transb, N, K)
    ref = alpha * a @ b + beta * C[:M].astype(np.float64)
    scale = np.abs(alpha) * np.abs(a) @ np.abs(b) + abs(beta) * np.abs(C[:M])
    tol = 1e-13 if dtype == np.float64 else 2e-5 * max(1.0, K / 700.0)
    assert np.all(np.abs(out[:M] - ref) <= tol * (scale + 1e-30))
    # rows beyond M (padding of ldc) untouched
    assert np.array_equal(out[M:], C[M:])


@pytest.mark.parametrize("M,N,K,transb", [(4133, 3900, 300, 1), (4133, 2100, 301, 0),
                                          (4133, 2100, 2100, 2), (300, 129, 203, 1),
                                          (300, 129, 203, 0), (200, 77, 77, 2)])
def test_gemm_reads_only_the_operands(M, N, K, transb):
    """NaN in every element the product must not read -- the leading-dimension padding of A
    and B, and for symmetric B its strict upper triangle -- leaves C bit-identical to the run
    on clean operands: row/column edges, the K tail of the last slice and the symmetric
    diagonal blocks stay inside the operands."""
    print(f"[gemm-nan] M={M} N={N} K={K} transb={transb}", flush=True)
    rng = np.random.default_rng(11 + transb)
    lda, ldb, ldc = M + 3, (N if transb == 1 else K) + 5, M
    A = np.asfortranarray(rng.standard_normal((lda, K)))
    B = np.asfortranarray(rng.standard_normal((ldb, N if transb == 0 else K)))
    C = np.asfortranarray(rng.standard_normal((ldc, N)))
    An, Bn = A.copy(order="F"), B.copy(order="F")
    An[M:] = np.nan
    Bn[(N if transb == 1 else K):] = np.nan
    if transb == 2:
        Bn[np.triu_indices(K, 1)] = np.nan
    outs = []
    for a_, b_ in ((A, B), (An, Bn)):
        dA, dB, dC = DeviceArray.from_host(a_), DeviceArray.from_host(b_), DeviceArray.from_host(C)
        dgemm_dev(transb, M, N, K, 0.5, dA.ptr, lda, dB.ptr, ldb, 1.0, dC.ptr, ldc)
        outs.append(dC.to_host())
    assert np.all(np.isfinite(outs[1]))
    assert np.array_equal(outs[0], outs[1])
    ref = 0.5 * A[:M] @ op_b(B, transb, N, K) + C
    assert np.allclose(outs[0], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


SPLIT_CASES = [(128, 128, 32, 1, 0.0), (130, 77, 33, 1, 1.0), (300, 129, 200, 0, -1.0),
               (1, 1, 1, 0, 0.5), (257, 383, 700, 1, 1.0), (64, 500, 5, 0, 0.0),
               (4133, 3900, 300, 1, 1.0), (4133, 3900, 301, 0, -1.0), (4133, 2100, 3001, 1, 1.0)]


@pytest.mark.parametrize("M,N,K,transb,beta", SPLIT_CASES)
def test_split_bf16_gemm_matches_numpy(M, N, K, transb, beta):
    """The split-BF16 product (hi*hi + hi*lo + lo*hi on BF16 MFMA, FP32 accumulation): each
    product carries a relative error of at most ~3 * 2^-16 (the dropped lo*lo term and the
    residue of the two-term split), so |C - C_ref| <= 1e-4 (|A||op(B)| + |beta C|) elementwise
    plus the FP32 accumulation bound; and the relative Frobenius error of the product stays
    below 1e-4, where one BF16 term per operand would give ~2e-3.  NaN planted in the
    leading-dimension padding must not reach C."""
    from moao_amd.reconstructor import sgemm_bf16x3_dev
    print(f"[gemm-bf16x3] M={M} N={N} K={K} transb={transb} beta={beta}", flush=True)
    rng = np.random.default_rng(M * 5 + N * 3 + K)
    lda, ldb, ldc = M + 3, (N if transb == 1 else K) + 5, M + 1
    A = np.asfortranarray(rng.standard_normal((lda, K)).astype(np.float32))
    B = np.asfortranarray(rng.standard_normal((ldb, N if transb == 0 else K)).astype(np.float32))
    C = np.asfortranarray(rng.standard_normal((ldc, N)).astype(np.float32))
    A[M:] = np.nan
    B[(N if transb == 1 else K):] = np.nan
    alpha = -0.75
    dA, dB, dC = DeviceArray.from_host(A), DeviceArray.from_host(B), DeviceArray.from_host(C)
    sgemm_bf16x3_dev(transb, M, N, K, alpha, dA.ptr, lda, dB.ptr, ldb, beta, dC.ptr, ldc)
    out = dC.to_host()
    a = A[:M].astype(np.float64)
    b = op_b(B, transb, N, K)
    prod = alpha * a @ b
    ref = prod + beta * C[:M].astype(np.float64)
    scale = np.abs(alpha) * np.abs(a) @ np.abs(b) + abs(beta) * np.abs(C[:M])
    tol = 1e-4 + 2e-5 * max(1.0, K / 700.0)
    err = np.abs(out[:M] - ref)
    assert np.all(np.isfinite(out[:M]))
    assert np.all(err <= tol * (scale + 1e-30)), float(np.max(err / (scale + 1e-30)))
    rel = np.linalg.norm(out[:M] - ref) / np.linalg.norm(prod)
    print(f"    max normalised error {np.max(err / (scale + 1e-30)):.2e}, "
          f"relative Frobenius {rel:.2e}", flush=True)
    assert rel <= 1e-4
    assert np.array_equal(out[M:], C[M:])
