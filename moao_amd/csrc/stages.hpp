// Device-side stage entry points shared by the C-ABI wrappers and the serving engine.  Each
// takes its stream and the scratch it may use explicitly, so independent callers (one per
// serving slot) never share a buffer.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace moao {

// matcov of one tile into device memory (MOAO_COMMON::matcov_tile semantics).  tables: the
// caller's upload buffer for the per-call tables; nullptr = the library context's buffer, used
// under the context lock.  Returns after the kernel has run (the tables may be reused).
int matcov_tile_launch(double* d_data, int nrows, int ncols, int xoffset, int yoffset, int lda,
                       const moao_tomo_view* t, int type_mat, hipStream_t stream, DevBuf* tables);

// Bytes of scratch the mixed-precision reconstructor needs for this problem.
size_t mixed_workspace_bytes(long n, long m, const moao_tomo_view* lowrank, int mode);

// Mixed-precision reconstructor (FP32 factorization of C0 = Cmm - lgs_cst U U^T, Woodbury,
// FP64 refinement to tol).  Cmm is read only; Ctm is overwritten by R.  ws: device scratch of
// at least mixed_workspace_bytes(n, m, lowrank) bytes, owned by the caller.  mode:
// MOAO_MIXED_FP32 (update products on FP32 MFMA) or MOAO_MIXED_BF16 (on the split-BF16 kernel).
int reconstructor_mixed_dev(long n, long m, const double* Cmm, long ldCmm, double* Ctm, long ldCtm,
                            const moao_tomo_view* lowrank, int mode, double tol, int max_iter,
                            double* resid_out, int* iters_out, void* ws, size_t ws_bytes,
                            hipStream_t s);

}  // namespace moao
