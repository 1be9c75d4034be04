// MFMA GEMM engine shared by the reconstructor and Cee/Cvv stages.
//
//   C_g[M_g x N] = alpha * A_g[M_g x K] * op(B) + beta * C_g       for each segment g
//
// Column-major, explicit leading dimensions.  op(B) = B^T with B stored N x K (transb = 1),
// op(B) = B with B stored K x N (transb = 0), or op(B) = B for a symmetric K x K matrix of which
// only the lower triangle is read (transb = 2, the B operand of dsymm(Right, Lower)).
// Up to kMaxSeg row segments share one B and one
// launch (grid.z = segment): the reconstructor uses this to update the Cmm rows and the Ctm
// rows of every target with a single launch.  A segment may bring its own B of the same shape
// (GemmSeg::B: a batch of independent products in one launch, e.g. one per target).  Segment 0
// (every segment with tri_all) may be restricted to its lower trapezoid (row + tri_off >= col):
// tiles wholly above that line are skipped and elements above it are not written.
//
// In-place use: when C_g == A_g and N <= the kernel's N tile (one column tile), each workgroup
// reads its whole row block of A before it writes C, and no other workgroup touches those rows,
// so C := A * op(B) in place is well defined (used for the triangular-inverse panel solves).
#pragma once

#include <hip/hip_runtime.h>

namespace moao {

constexpr int kMaxSeg = 8;

// Split-BF16 planes of an FP32 matrix region: the hi term of element (row, k) at
// hi[row * ld + k], the lo term at hi[ps + row * ld + k] (k contiguous within a row; ld a
// multiple of 8 and at least the row's last chunk of 32 k).
struct Planes {
    const unsigned short* hi;
    long ld;
    long ps;
};

template <typename T>
struct GemmSeg {
    const T* A;
    T* C;
    long lda, ldc;
    int M;
    Planes ap;  // split = 2: A's planes (row = row of this segment)
    const T* B;  // optional: this segment's own B (same shape and ldb), else GemmArgs::B
};

template <typename T>
struct GemmArgs {
    GemmSeg<T> seg[kMaxSeg];
    int nseg;
    const T* B;
    long ldb;
    int N, K;
    int transb;
    int tri;      // segment 0 lower-trapezoid restriction
    int tri_off;
    int tri_all;  // with tri: every segment restricted, not only segment 0
    T alpha, beta;
    // transb = 2 only, optional: full symmetric copies of B's 128 x 128 diagonal blocks
    // (gemm_symdiag); with it the launch runs on the direct-to-LDS kernel.
    const T* symdiag;
    // FP32 only: 2 = run the product on BF16 MFMA from split-BF16 planes of the operands
    // (hi = bf16(a), lo = bf16(a - hi); hi*hi + hi*lo + lo*hi accumulated in FP32; A from
    // seg[].ap, op(B) element (k, n) from bp row n); 0 = FP32 MFMA.
    int split;
    Planes bp;  // split = 2: B's planes (row = column n of op(B))
    // Optional device status word: when it is non-zero at launch time the whole launch is a
    // no-op.  The factorization points it at its dpotrf info, so the updates and solves queued
    // after a failing pivot do nothing and no host synchronisation is needed to stop them.
    const int* skip_if;
};

// Elements of the gemm_symdiag copy of a K x K matrix.
size_t gemm_symdiag_elems(long K);
// out[b] (128 x 128, leading dimension 128) = the full symmetric diagonal block b of the K x K
// matrix whose lower triangle B holds; what a transb = 2 launch reads where its K slices cross
// the diagonal.
int gemm_symdiag(const double* B, long ldb, int K, double* out, hipStream_t stream);

// Planes of an FP32 column-major region (rows x cols at A, leading dimension lda):
// transposed = 1 writes element (r, c) as plane row r, k = c (hi[r * ld + c]); transposed = 0 as
// plane row c, k = r (hi[c * ld + r]).
int split_planes(const float* A, long lda, int rows, int cols, unsigned short* hi, long ld,
                 long ps, int transposed, hipStream_t stream);

// Launches on `stream`; returns a MOAO_* status (argument checks only, launch is async).
int gemm_f64(const GemmArgs<double>& a, hipStream_t stream);
int gemm_f32(const GemmArgs<float>& a, hipStream_t stream);

inline int gemm(const GemmArgs<double>& a, hipStream_t s) { return gemm_f64(a, s); }
inline int gemm(const GemmArgs<float>& a, hipStream_t s) { return gemm_f32(a, s); }

// Convenience for a single segment.
template <typename T>
inline GemmArgs<T> gemm_args(int M, int N, int K, T alpha, const T* A, long lda, const T* B,
                             long ldb, int transb, T beta, T* C, long ldc) {
    GemmArgs<T> g{};
    g.seg[0] = {A, C, lda, ldc, M};
    g.nseg = 1;
    g.B = B;
    g.ldb = ldb;
    g.N = N;
    g.K = K;
    g.transb = transb;
    g.tri = 0;
    g.tri_off = 0;
    g.tri_all = 0;
    g.alpha = alpha;
    g.beta = beta;
    g.symdiag = nullptr;
    g.split = 0;
    g.bp = Planes{nullptr, 0, 0};
    g.skip_if = nullptr;
    return g;
}

}  // namespace moao
