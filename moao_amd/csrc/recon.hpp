// Reconstructor R = Ctm * Cmm^{-1} on the device (MOAO_LAPACK::reconstructor semantics).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "gemm.hpp"
#include "leaf.hpp"

namespace moao {

// Split-BF16 planes of the mixed BF16 solve (FP32 storage): lp holds L (n x n) with plane row
// r = row of L and k = column, lq holds L with plane row = column of L and k = row, xp the
// (single) row group with plane row = group row and k = column; all with leading dimension ld
// (a multiple of 64, >= n), hi terms at p[...], lo terms at p[ps + ...].
struct SplitWs {
    unsigned short* lp;
    unsigned short* lq;
    unsigned short* xp;
    long ld, lps, lqs, xps;
    // operand views: rows from r0, k from c0
    Planes l_rows(long r0, long c0) const { return Planes{lp + r0 * ld + c0, ld, lps}; }
    Planes l_cols(long c0, long r0) const { return Planes{lq + c0 * ld + r0, ld, lqs}; }
    Planes x_rows(long c0) const { return Planes{xp + c0, ld, xps}; }
};

// A block of rows that rides along the factorization as extra rows of the tall matrix
// [Cmm; group_0; group_1; ...]: after potrf_tall each group holds X = B * L^{-T}.
template <typename T>
struct RowGroup {
    T* p;
    long ld;
    int rows;
};

// Recursive factorization of the n x n SPD block (lower, in place) with the row groups
// solved by L^T on the way (first right-side solve fused).  invL receives the inverse of every
// kLeaf leaf (ceil(n/kLeaf)*kLeaf*kLeaf elements); *info the dpotrf status (device int).
// sw (FP32 storage only): run the update products (not the leaves' panel products) on the
// split-BF16 planes kernel, writing the planes of every finished leaf column into sw->lp and
// sw->xp (one row group at most).
template <typename T>
int potrf_tall(long n, T* A, long lda, const std::vector<RowGroup<T>>& groups, T* invL,
               int* info, hipStream_t s, const SplitWs* sw = nullptr);
// Y * L = X for every group (second right-side solve), with L / invL from potrf_tall.
// skip: optional device status word; every launch is a no-op once it is non-zero.
template <typename T>
int trsm_right_lower_notrans(long n, const T* A, long lda, const T* invL,
                             const std::vector<RowGroup<T>>& groups, hipStream_t s,
                             const int* skip = nullptr, const SplitWs* sw = nullptr);
// X * L^T = B for every group (first right-side solve, standalone).
template <typename T>
int trsm_right_lower_trans(long n, const T* A, long lda, const T* invL,
                           const std::vector<RowGroup<T>>& groups, hipStream_t s,
                           const int* skip = nullptr, const SplitWs* sw = nullptr);

// FP64 path, stream-ordered (no host synchronisation).  dCmm (n x n, lower) is overwritten by
// its Cholesky factor L; each of the ngal target blocks Ctm_g = dCtm + g*n*m (m x n, leading
// dimension ldCtm) is overwritten by R_g.  *dinfo (device) receives the dpotrf status: 0, or
// the 1-based order of the first non-positive pivot, after which every later launch of the
// solve is a no-op (the Ctm blocks then hold partial results).  dinvL: workspace of
// recon_invl_doubles(n) doubles.
int reconstructor_f64_dev(long n, long m, double* dCmm, long ldCmm, double* dCtm, long ldCtm,
                          long ngal, double* dinvL, int* dinfo, hipStream_t stream);

// The same on FP32 storage (FP32 MFMA updates, FP64 leaves); dinvL holds
// recon_invl_doubles(n) floats.
int reconstructor_f32_dev(long n, long m, float* dCmm, long ldCmm, float* dCtm, long ldCtm,
                          long ngal, float* dinvL, int* dinfo, hipStream_t stream);

// Takes the current device's address of the allocation-free entry's static scratch (once per
// device, with that device current); called by the library's init.
int noalloc_scratch_setup();

// Number of elements of the inverse-leaf workspace.
size_t recon_invl_doubles(long n);
// Bytes of one FP64 solve's workspace: the inverse leaves, then (last 256 bytes) the status.
size_t recon_workspace_bytes(long n);

}  // namespace moao
