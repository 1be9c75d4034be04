// Reconstructor R = Ctm * Cmm^{-1} on the device (MOAO_LAPACK::reconstructor semantics).
#pragma once

#include <hip/hip_runtime.h>

namespace moao {

// FP64 path.  dCmm (n x n, lower) is overwritten by its Cholesky factor L; each of the ngal
// target blocks Ctm_g = dCtm + g*n*m (m x n, leading dimension ldCtm) is overwritten by R_g.
// On a non-positive pivot *info = its 1-based order and the Ctm blocks hold partial results.
// dinvL: workspace of ceil(n/kLeaf)*kLeaf*kLeaf doubles; dinfo: one device int.
int reconstructor_f64_dev(long n, long m, double* dCmm, long ldCmm, double* dCtm, long ldCtm,
                          long ngal, double* dinvL, int* dinfo, hipStream_t stream);

// Size in doubles of the dinvL workspace.
size_t recon_invl_doubles(long n);

}  // namespace moao
