// Diagonal-leaf kernels of the reconstructor: FP64 Cholesky of one nb x nb (nb <= 128) block
// held entirely in LDS, followed by the in-place inverse of its triangular factor.
//
// LAPACK semantics of dpotrf('L') for the block: the lower triangle is overwritten by L, the
// strictly-upper part is not touched; a non-positive (or NaN) pivot at local column j stops
// the factorization and records info = j_global + 1 (first failure wins; later leaves see the
// flag and return at once).  The inverse L^{-1} (lower, zero above the diagonal) goes to a
// separate nb x nb buffer (leading dimension kLeaf) and turns every panel triangular solve
// into an MFMA GEMM.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "leaf.hpp"
#include "profile.hpp"

namespace moao {
namespace {

constexpr int NTL = 256;  // lanes per leaf workgroup
constexpr int LD = kLeaf + 1;

// S is column-major in LDS: S[c * LD + r] = A(r, c).
__global__ void __launch_bounds__(NTL) leaf_potrf_inv(double* A, long lda, int nb, double* inv,
                                                      int* info, int j0) {
    __shared__ double S[kLeaf * LD];
    __shared__ double x[kLeaf];
    __shared__ int fail;
    if (*(volatile int*)info != 0) return;  // an earlier leaf failed: R is not computed
    const int tid = threadIdx.x;
    if (tid == 0) fail = 0;
    for (int e = tid; e < nb * nb; e += NTL) {
        const int r = e % nb, c = e / nb;
        S[c * LD + r] = (r >= c) ? A[r + (long)c * lda] : 0.0;
    }
    __syncthreads();

    // Right-looking unblocked Cholesky.  Lane (row i, half h) updates columns k = j+1+h, +2, ...
    const int i = tid % kLeaf, h = tid / kLeaf;
    for (int j = 0; j < nb; ++j) {
        const double d = S[j * LD + j];
        if (!(d > 0.0)) {  // uniform across the workgroup
            if (tid == 0) atomicCAS(info, 0, j0 + j + 1);
            return;
        }
        const double sd = sqrt(d);
        __syncthreads();  // everyone has read S[j][j]
        if (i < nb && i > j && h == 0) S[j * LD + i] /= sd;
        if (tid == 0) S[j * LD + j] = sd;
        __syncthreads();
        if (i < nb && i > j) {
            const double lij = S[j * LD + i];
            for (int k = j + 1 + h; k <= i; k += 2) S[k * LD + i] -= lij * S[j * LD + k];
        }
        __syncthreads();
    }
    // L back to global (lower part only).
    for (int e = tid; e < nb * nb; e += NTL) {
        const int r = e % nb, c = e / nb;
        if (r >= c) A[r + (long)c * lda] = S[c * LD + r];
    }
    // In-place inverse of the lower-triangular S, columns from last to first:
    //   S[j][j] <- 1/L_jj;  S[j+1:, j] <- -(inv(L)[j+1:, j+1:] * L[j+1:, j]) / L_jj
    for (int j = nb - 1; j >= 0; --j) {
        __syncthreads();
        if (tid < nb && tid > j) x[tid] = S[j * LD + tid];
        const double ajj = 1.0 / S[j * LD + j];  // read by all before lane 0 overwrites it
        __syncthreads();
        if (tid < nb && tid > j) {
            double acc = 0.0;
            for (int k = j + 1; k <= tid; ++k) acc += S[k * LD + tid] * x[k];
            S[j * LD + tid] = -acc * ajj;
        }
        if (tid == 0) S[j * LD + j] = ajj;
    }
    __syncthreads();
    for (int e = tid; e < kLeaf * kLeaf; e += NTL) {
        const int r = e % kLeaf, c = e / kLeaf;
        inv[r + (long)c * kLeaf] = (r < nb && c < nb && r >= c) ? S[c * LD + r] : 0.0;
    }
}

}  // namespace

int leaf_potrf_inv_launch(double* A, long lda, int nb, double* inv, int* info, int j0,
                          hipStream_t stream) {
    if (nb < 1 || nb > kLeaf) {
        set_error("leaf_potrf: bad block size");
        return MOAO_ERR_ARG;
    }
    ProfScope ps(kProfLeaf, (double)nb * nb * nb / 3.0, stream);
    hipLaunchKernelGGL(leaf_potrf_inv, dim3(1), dim3(NTL), 0, stream, A, lda, nb, inv, info, j0);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

}  // namespace moao
