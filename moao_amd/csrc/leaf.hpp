// Diagonal-leaf Cholesky + triangular inverse (see leaf.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace moao {

constexpr int kLeaf = 128;  // leaf block size of the recursive factorization

// Factor the nb x nb block at A (lower, in place) and write inv(L) to inv (kLeaf x kLeaf).
// Storage type T (double or float); the factorization itself always runs in FP64 in LDS.
template <typename T>
int leaf_potrf_inv_launch(T* A, long lda, int nb, T* inv, int* info, int j0, hipStream_t stream);

}  // namespace moao
