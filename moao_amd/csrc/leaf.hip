// Diagonal-leaf kernels of the reconstructor: FP64 Cholesky of one nb x nb (nb <= 128) block
// held entirely in LDS, followed by the in-place inverse of its triangular factor.
//
// LAPACK semantics of dpotrf('L') for the block: the lower triangle is overwritten by L, the
// strictly-upper part is not touched; a non-positive (or NaN) pivot at local column j stops
// the factorization and records info = j_global + 1 (first failure wins; later leaves see the
// flag and return at once).  The inverse L^{-1} (lower, zero above the diagonal) goes to a
// separate kLeaf x kLeaf buffer and turns every panel triangular solve into an MFMA GEMM.
//
// Both halves are blocked by kB = 16 columns so a leaf needs ~6 workgroup barriers per block
// instead of ~3 per column:
//   factor: [wave 0] unblocked Cholesky of the 16x16 diagonal block (wave-synchronous LDS);
//           [all] rows below solve against it by substitution (one lane per row);
//           [all] 2x2-register-blocked rank-16 update of the trailing lower triangle.
//   invert: block columns from last to first: [wave 0] D = inv(L_jj) by substitution;
//           [all] T = inv(L)[below, below] * L[below, j];  [all] column j of inv(L) = -T * D.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "leaf.hpp"
#include "profile.hpp"

namespace moao {
namespace {

constexpr int NTL = 256;       // lanes per leaf workgroup
constexpr int LD = kLeaf + 1;  // odd stride: a row across columns hits distinct banks
constexpr int kB = 16;         // inner block

#define S_(r, c) S[(c) * LD + (r)]

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <typename TS>
__global__ void __launch_bounds__(NTL) leaf_potrf_inv(TS* A, long lda, int nb, TS* inv,
                                                      int* info, int j0) {
    __shared__ double S[kLeaf * LD];
    __shared__ double T[(kLeaf - kB) * kB];  // inverse: product of the trailing part
    __shared__ double D[kB * kB];            // inverse of one diagonal block
    __shared__ int fail;
    if (*(volatile int*)info != 0) return;  // an earlier leaf failed: R is not computed
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) fail = 0;
    // a wave per column, lanes down the rows (no runtime division in the index math)
    for (int c = wave; c < nb; c += NTL / 64)
        for (int r = lane; r < nb; r += 64) S_(r, c) = (r >= c) ? (double)A[r + (long)c * lda] : 0.0;
    __syncthreads();

    // ------------------------------------------------------------------ factorization
    for (int kb = 0; kb < nb; kb += kB) {
        const int bs = min(kB, nb - kb);
        if (wave == 0) {
            const int i = lane & 15, kq = lane >> 4;
            for (int j = 0; j < bs; ++j) {
                const double d = S_(kb + j, kb + j);
                if (!(d > 0.0)) {  // same value in every lane
                    if (lane == 0) {
                        fail = 1;
                        atomicCAS(info, 0, j0 + kb + j + 1);
                    }
                    break;
                }
                const double sd = sqrt(d);
                wave_sync();
                if (lane > j && lane < bs) S_(kb + lane, kb + j) /= sd;
                if (lane == 0) S_(kb + j, kb + j) = sd;
                wave_sync();
                if (i > j && i < bs) {
                    const double lij = S_(kb + i, kb + j);
                    for (int k = j + 1 + kq; k <= i; k += 4)
                        S_(kb + i, kb + k) -= lij * S_(kb + k, kb + j);
                }
                wave_sync();
            }
        }
        __syncthreads();
        if (fail) return;
        // rows below the diagonal block: x * L_bb^T = a  (forward substitution per row)
        const int rest = nb - kb - bs;
        if (tid < rest) {
            const int r = kb + bs + tid;
            double x[kB];
#pragma unroll
            for (int c = 0; c < kB; ++c) x[c] = c < bs ? S_(r, kb + c) : 0.0;
#pragma unroll
            for (int c = 0; c < kB; ++c) {
                if (c < bs) {
                    double v = x[c];
#pragma unroll
                    for (int k = 0; k < c; ++k) v -= x[k] * S_(kb + c, kb + k);
                    x[c] = v / S_(kb + c, kb + c);
                }
            }
#pragma unroll
            for (int c = 0; c < kB; ++c)
                if (c < bs) S_(r, kb + c) = x[c];
        }
        __syncthreads();
        // trailing update of the lower triangle, 2x2 register blocks
        if (rest > 0) {
            const int h = (rest + 1) / 2;
            const int nblk = h * (h + 1) / 2;
            const int base = kb + bs;
            for (int blk = tid; blk < nblk; blk += NTL) {
                int bi = (int)((sqrtf(8.0f * blk + 1.0f) - 1.0f) * 0.5f);
                while ((bi + 1) * (bi + 2) / 2 <= blk) ++bi;
                while (bi * (bi + 1) / 2 > blk) --bi;
                const int bj = blk - bi * (bi + 1) / 2;
                const int r0 = base + 2 * bi, c0 = base + 2 * bj;
                const bool r1ok = r0 + 1 < nb, c1ok = c0 + 1 < nb;
                double a00 = 0, a01 = 0, a10 = 0, a11 = 0;
                for (int k = 0; k < bs; ++k) {
                    const double p0 = S_(r0, kb + k), p1 = r1ok ? S_(r0 + 1, kb + k) : 0.0;
                    const double q0 = S_(c0, kb + k), q1 = c1ok ? S_(c0 + 1, kb + k) : 0.0;
                    a00 += p0 * q0;
                    a01 += p0 * q1;
                    a10 += p1 * q0;
                    a11 += p1 * q1;
                }
                S_(r0, c0) -= a00;
                if (c1ok && r0 >= c0 + 1) S_(r0, c0 + 1) -= a01;
                if (r1ok) S_(r0 + 1, c0) -= a10;
                if (r1ok && c1ok) S_(r0 + 1, c0 + 1) -= a11;
            }
        }
        __syncthreads();
    }
    // L back to global (lower part only).
    for (int c = wave; c < nb; c += NTL / 64)
        for (int r = c + lane; r < nb; r += 64) A[r + (long)c * lda] = (TS)S_(r, c);

    // ------------------------------------------------------------------ inverse
    const int nblocks = (nb + kB - 1) / kB;
    for (int jb = nblocks - 1; jb >= 0; --jb) {
        const int c0 = jb * kB, bs = min(kB, nb - c0);
        const int below = nb - c0 - bs;
        // D = inv(L_jj): lane c computes column c by forward substitution
        if (wave == 0 && lane < bs) {
            const int c = lane;
            double x[kB];
#pragma unroll
            for (int r = 0; r < kB; ++r) {
                double v = 0.0;
                if (r >= c && r < bs) {
                    v = (r == c) ? 1.0 : 0.0;
#pragma unroll
                    for (int k = 0; k < r; ++k)
                        if (k >= c) v -= S_(c0 + r, c0 + k) * x[k];
                    v /= S_(c0 + r, c0 + r);
                }
                x[r] = v;
            }
#pragma unroll
            for (int r = 0; r < kB; ++r) D[r * kB + c] = x[r];  // D[r][c]
        }
        // T = inv(L)[below, below] * L[below, c0:c0+bs]   (inv(L) trailing part already in S)
        // thread -> (row rr = tid/16 + 16 i, column c = tid%16): no runtime division
        const int tc = tid & (kB - 1), tr = tid / kB;
        if (tc < bs)
            for (int rr = tr; rr < below; rr += NTL / kB) {
                const int r = c0 + bs + rr;
                double v = 0.0;
                for (int k = c0 + bs; k <= r; ++k) v += S_(r, k) * S_(k, c0 + tc);
                T[tc * (kLeaf - kB) + rr] = v;
            }
        __syncthreads();
        // column block jb of inv(L): rows below = -T * D; diagonal block = D
        if (tc < bs) {
            for (int rr = tr; rr < below; rr += NTL / kB) {
                double v = 0.0;
                for (int k = tc; k < bs; ++k) v += T[k * (kLeaf - kB) + rr] * D[k * kB + tc];
                S_(c0 + bs + rr, c0 + tc) = -v;
            }
            if (tr < bs) S_(c0 + tr, c0 + tc) = tr >= tc ? D[tr * kB + tc] : 0.0;
        }
        __syncthreads();
    }
    for (int c = wave; c < kLeaf; c += NTL / 64)
        for (int r = lane; r < kLeaf; r += 64)
            inv[r + (long)c * kLeaf] = (TS)((r < nb && c < nb && r >= c) ? S_(r, c) : 0.0);
}

#undef S_

}  // namespace

template <typename T>
int leaf_potrf_inv_launch(T* A, long lda, int nb, T* inv, int* info, int j0, hipStream_t stream) {
    if (nb < 1 || nb > kLeaf) {
        set_error("leaf_potrf: bad block size");
        return MOAO_ERR_ARG;
    }
    ProfScope ps(kProfLeaf, (double)nb * nb * nb / 3.0, stream);
    hipLaunchKernelGGL(leaf_potrf_inv<T>, dim3(1), dim3(NTL), 0, stream, A, lda, nb, inv, info,
                       j0);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

template int leaf_potrf_inv_launch<double>(double*, long, int, double*, int*, int, hipStream_t);
template int leaf_potrf_inv_launch<float>(float*, long, int, float*, int*, int, hipStream_t);

}  // namespace moao
