// Diagonal-leaf kernels of the reconstructor: FP64 Cholesky of one nb x nb (nb <= 128) block
// held entirely in LDS, followed by the in-place inverse of its triangular factor.
//
// LAPACK semantics of dpotrf('L') for the block: the lower triangle is overwritten by L, the
// strictly-upper part is not touched; a non-positive (or NaN) pivot at local column j stops
// the factorization and records info = j_global + 1 (first failure wins; later leaves see the
// flag and return at once); the leaf's 16-column blocks before the failing pivot's block are
// stored, the rest of the block keeps its input.  The inverse L^{-1} (lower, zero above the diagonal) goes to a
// separate kLeaf x kLeaf buffer and turns every panel triangular solve into an MFMA GEMM.
//
// One workgroup (4 waves) per leaf; the block is zero-padded to kLeaf x kLeaf in LDS and
// processed by 16-column blocks:
//   load:   batched global loads (16 in flight per lane) before the LDS stores;
//   factor: [wave 0] 16x16 diagonal block in registers (readlane broadcasts, 1/sqrt of each
//           pivot by a refined hardware estimate), then D = inv(L_bb) by right-looking
//           substitution, kept in the unused block above the diagonal; [all] rows below as
//           A * D^T on f64 MFMA; the rank-16 update of the trailing lower triangle (f64 MFMA)
//           runs one block late, beside the next diagonal factor: wave 0 updates that
//           diagonal block first, waves 1-3 the rest;
//   invert: block columns from last to first: [all] T = inv(L)[below, below] * L[below, j]
//           and column j of inv(L) = -T * D, both on f64 MFMA, 16-row tiles per wave;
//   store:  L and inv(L) with 16 LDS reads per lane in flight before the global stores.
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

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Value of v in lane l (l wave-uniform).
__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// 1/sqrt(d) for d > 0: hardware estimate refined by two Newton steps (no division, no
// correctly-rounded sqrt sequence on the pivot's critical path).
__device__ __forceinline__ double rsqrt_nr(double d) {
    double r = __builtin_amdgcn_rsq(d);
    r = r * (1.5 - 0.5 * d * r * r);
    r = r * (1.5 - 0.5 * d * r * r);
    return r;
}

// 16 x 16 block product on MFMA: acc += X[r0:r0+16, k0:k0+16*nc] * Y[k0:k0+16*nc, c0:c0+16]
// (nc <= MAXC chunks of 16) with X element (r, k) at X[r + k*ldx] and Y element (k, c) at
// Y[yk*k + yc*c] (LDS).  Lane l supplies X[r0 + l%16][k] and Y[k][c0 + l%16] for
// k = k0 + 4s + l/16; accumulator item t of lane l is row l/16 + 4t, column l%16.  The loop is
// unrolled and the loads of chunk q+1 are issued before the four MFMAs of chunk q, so the LDS
// latency is paid once per call rather than per MFMA; the accumulation order is k ascending.
template <int MAXC>
__device__ __forceinline__ f64x4 mfma_block(const double* X, int ldx, int r0, const double* Y,
                                             int yk, int yc, int c0, int k0, int nc, f64x4 acc,
                                             int lane) {
    const int li = lane & 15, lk = lane >> 4;
    const double* xp = X + (r0 + li) + (k0 + lk) * ldx;  // element (r0 + li, k0 + lk)
    const double* yp = Y + yk * (k0 + lk) + yc * (c0 + li);
    double a[2][4], b[2][4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        a[0][s] = xp[4 * s * ldx];
        b[0][s] = yp[4 * s * yk];
    }
#pragma unroll
    for (int q = 0; q < MAXC; ++q) {
        if (q >= nc) break;
        const int cur = q & 1;
        if (q + 1 < nc) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                a[cur ^ 1][s] = xp[(16 * (q + 1) + 4 * s) * ldx];
                b[cur ^ 1][s] = yp[(16 * (q + 1) + 4 * s) * yk];
            }
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[cur][s], b[cur][s], acc, 0, 0, 0);
    }
    return acc;
}

template <typename TS>
__global__ void __launch_bounds__(NTL) leaf_potrf_inv(TS* A, long lda, int nb, TS* inv,
                                                      int* info, int j0) {
    __shared__ double S[kLeaf * LD];
    __shared__ double T[(kLeaf - kB) * kB];  // inverse: product of the trailing part
    __shared__ double D[kB * kB];            // inverse of one diagonal block
    __shared__ double rdiag[kLeaf];          // 1 / L_jj
    __shared__ int fail;
    if (*(volatile int*)info != 0) return;  // an earlier leaf failed: R is not computed
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar branches
    if (tid == 0) fail = 0;
    // Load the lower triangle, zero-padded to kLeaf x kLeaf.  A wave owns columns
    // wave, wave+4, ...; each batch issues 16 independent loads per lane before any LDS store
    // (addresses clamped into the block, values outside the lower triangle replaced by 0).
    {
        constexpr int CB = 8;
        const int rA = min(lane, nb - 1), rB = min(lane + 64, nb - 1);
        for (int cb = wave; cb < kLeaf; cb += 4 * CB) {
            double v0[CB], v1[CB];
#pragma unroll
            for (int q = 0; q < CB; ++q) {
                const long cc = min(cb + 4 * q, nb - 1);
                v0[q] = (double)A[rA + cc * lda];
                v1[q] = (double)A[rB + cc * lda];
            }
#pragma unroll
            for (int q = 0; q < CB; ++q) {
                const int c = cb + 4 * q;
                S_(lane, c) = (c < nb && lane < nb && lane >= c) ? v0[q] : 0.0;
                S_(lane + 64, c) = (c < nb && lane + 64 < nb && lane + 64 >= c) ? v1[q] : 0.0;
            }
        }
    }
    __syncthreads();

    // Columns [0, ncol) of L back to global (lower part only), 16 LDS reads per lane in flight
    // before the stores.
    auto store_L = [&](int ncol) {
        for (int cb = wave; cb < ncol; cb += 4 * 8) {
            double v0[8], v1[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int c = min(cb + 4 * q, kLeaf - 1);
                v0[q] = S_(lane, c);
                v1[q] = S_(lane + 64, c);
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int c = cb + 4 * q;
                if (c >= ncol) break;
                if (lane >= c && lane < nb) A[lane + (long)c * lda] = (TS)v0[q];
                if (lane + 64 >= c && lane + 64 < nb) A[lane + 64 + (long)c * lda] = (TS)v1[q];
            }
        }
    };

    // ------------------------------------------------------------------ factorization
    // Trailing update of block kp's trailing lower triangle, tile t (16 x 16, ti >= tj, tiles
    // numbered row by row), K = 16 on MFMA: S -= P * P^T with P[r][k] = S_(r, kp + k).
    auto trail_tile = [&](int kp, int t) {
        const int base = kp + kB;
        int ti = 0, acc0 = 0;
        while (acc0 + ti + 1 <= t) acc0 += ++ti;
        const int tj = t - acc0;
        f64x4 acc = {0, 0, 0, 0};
        // B operand is P^T: element (k, c) = S_(c, kp + k)
        acc = mfma_block<1>(&S_(0, kp), LD, base + 16 * ti, &S_(0, kp), LD, 1, base + 16 * tj, 0,
                            1, acc, lane);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int r = base + 16 * ti + (lane >> 4) + 4 * e;
            const int c = base + 16 * tj + (lane & 15);
            if (r < nb && c < nb && r >= c) S_(r, c) -= acc[e];
        }
    };
    for (int kb = 0; kb < nb; kb += kB) {
        const int bs = min(kB, nb - kb);
        // Look-ahead: the previous block's trailing update runs here, beside this block's
        // diagonal factor.  Wave 0 takes tile 0 (this block's diagonal block) and then factors
        // it; waves 1-3 take the other tiles, which this factor neither reads nor writes.
        if (kb > 0) {
            const int ht = (nb - kb + 15) / 16, ntile = ht * (ht + 1) / 2;
            if (wave == 0) trail_tile(kb - kB, 0);
            else
                for (int t = wave; t < ntile; t += NTL / 64 - 1) trail_tile(kb - kB, t);
        }
        if (wave == 0) {
            // Diagonal block: lane i holds row i in registers; column j is broadcast with
            // readlane (uniform lane index), so the block needs no LDS round trips.
            const int i = lane;
            double a[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) a[k] = (i < bs && k <= i) ? S_(kb + i, kb + k) : 0.0;
            bool bad = false;
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                if (j < bs && !bad) {
                    const double d = lane_value(a[j], j);
                    if (!(d > 0.0)) {  // same value in every lane
                        bad = true;
                        if (lane == 0) {
                            fail = 1;
                            atomicCAS(info, 0, j0 + kb + j + 1);
                        }
                    } else {
                        const double rsd = rsqrt_nr(d), sd = d * rsd;
                        a[j] = (i == j) ? sd : a[j] * rsd;
                        if (lane == 0) rdiag[kb + j] = rsd;
#pragma unroll
                        for (int k = j + 1; k < kB; ++k) {
                            const double lkj = lane_value(a[j], k);
                            if (i >= k) a[k] -= a[j] * lkj;
                        }
                    }
                }
            }
            if (i < bs)
#pragma unroll
                for (int k = 0; k < kB; ++k)
                    if (k <= i) S_(kb + i, kb + k) = a[k];
            // D = inv(L_bb) (same wave, so the LDS writes above are visible): lane c builds
            // column c right-looking from e_c; rows at or past bs are zero padding.  D is kept
            // in the unused block above the diagonal, S_(kb + r, kb + kB + c), or in D[] for
            // the last block; the rows below and the inverse phase both use it.
            if (lane < kB) {
                const int c = lane;
                double x[kB];
#pragma unroll
                for (int r = 0; r < kB; ++r) x[r] = (r == c) ? 1.0 : 0.0;
#pragma unroll
                for (int r = 0; r < kB; ++r) {
                    x[r] = r < bs ? x[r] * rdiag[kb + r] : 0.0;
#pragma unroll
                    for (int r2 = r + 1; r2 < kB; ++r2) x[r2] -= S_(kb + r2, kb + r) * x[r];
                }
                if (kb + kB < kLeaf) {
#pragma unroll
                    for (int r = 0; r < kB; ++r) S_(kb + r, kb + kB + c) = x[r];
                } else {
#pragma unroll
                    for (int r = 0; r < kB; ++r) D[r * kB + c] = x[r];
                }
            }
        }
        __syncthreads();
        if (fail) {
            // the 16-column blocks before the failing one are final: store them (their rows
            // below included), leave the failing block and the rest as they were
            store_L(kb);
            return;
        }
        // rows below the diagonal block: L[r, kb:kb+16] = A[r, kb:kb+16] * D^T on MFMA, a
        // 16-row tile per wave turn, in place (bs is 16 whenever rows remain below the block)
        const int rest = nb - kb - bs;
        if (rest > 0) {
            const bool up = kb + kB < kLeaf;
            const double* Dp = up ? &S_(kb, kb + kB) : D;  // D element (r, c) at Dp[r*dr + c*dc]
            const int dr = up ? 1 : kB, dc = up ? LD : 1;
            const int ht = (rest + 15) / 16;
            for (int i = wave; i < ht; i += NTL / 64) {
                const int r0 = kb + kB + 16 * i;
                f64x4 acc = {0, 0, 0, 0};
                // Y element (k, c) = D^T(k, c) = D(c, k)
                acc = mfma_block<1>(&S_(0, kb), LD, r0, Dp, dc, dr, 0, 0, 1, acc, lane);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = r0 + (lane >> 4) + 4 * e;
                    if (r < nb) S_(r, kb + (lane & 15)) = acc[e];
                }
            }
        }
        __syncthreads();
    }
    store_L(nb);

    // ------------------------------------------------------------------ inverse
    const int nblocks = (nb + kB - 1) / kB;
    for (int jb = nblocks - 1; jb >= 0; --jb) {
        const int c0 = jb * kB, bs = min(kB, nb - c0);
        const int below = nb - c0 - bs;
        const int kbeg = c0 + bs;
        // D = inv(L_jj), computed during the factorization: copy it out of its storage above
        // the diagonal (the last block left it in D[] already); used after the next barrier
        if (c0 + kB < kLeaf)
            for (int t = tid; t < kB * kB; t += NTL) D[t] = S_(c0 + t / kB, c0 + kB + t % kB);
        // T = inv(L)[below, below] * L[below, c0:c0+16] on MFMA, a 16-row tile per wave turn;
        // inv(L) of the trailing part is lower triangular (zero above, zero padding below nb),
        // so row tile i needs k < 16 (i + 1) only.  Tiles go to the waves in snake order
        // (0 1 2 3, then 7 6 5 4), which evens out the longer products of the later tiles.
        const int nti = (below + 15) / 16;
        for (int r = 0; 4 * r < nti; ++r) {
            const int i = (r & 1) ? 4 * r + 3 - wave : 4 * r + wave;
            if (i >= nti) continue;
            f64x4 acc = {0, 0, 0, 0};
            acc = mfma_block<kLeaf / kB - 1>(&S_(0, kbeg), LD, 16 * i + kbeg, &S_(kbeg, c0), 1,
                                             LD, 0, 0, i + 1, acc, lane);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                T[(lane & 15) * (kLeaf - kB) + 16 * i + (lane >> 4) + 4 * e] = acc[e];
        }
        __syncthreads();
        // column block jb of inv(L): rows below = -T * D; diagonal block = D
        for (int i = wave; i < nti; i += NTL / 64) {
            f64x4 acc = {0, 0, 0, 0};
            acc = mfma_block<1>(T, kLeaf - kB, 16 * i, D, kB, 1, 0, 0, 1, acc, lane);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int rr = 16 * i + (lane >> 4) + 4 * e;
                if (rr < below) S_(kbeg + rr, c0 + (lane & 15)) = -acc[e];
            }
        }
        {
            const int tc = tid & (kB - 1), tr = tid / kB;
            if (tc < bs && tr < bs) S_(c0 + tr, c0 + tc) = tr >= tc ? D[tr * kB + tc] : 0.0;
        }
        __syncthreads();
    }
    for (int cb = wave; cb < kLeaf; cb += 4 * 8) {
        double v0[8], v1[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            v0[q] = S_(lane, cb + 4 * q);
            v1[q] = S_(lane + 64, cb + 4 * q);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = cb + 4 * q;
            inv[lane + (long)c * kLeaf] = (TS)((lane < nb && c < nb && lane >= c) ? v0[q] : 0.0);
            inv[lane + 64 + (long)c * kLeaf] =
                (TS)((lane + 64 < nb && c < nb && lane + 64 >= c) ? v1[q] : 0.0);
        }
    }
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
