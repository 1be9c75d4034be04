// Mixed-precision reconstructor: R = Ctm * Cmm^{-1} to a stated tolerance with the O(n^3)
// factorization and the O(m n^2) solves in low precision and the residuals in FP64.  Two modes:
// MOAO_MIXED_FP32 runs the updates on FP32 MFMA; MOAO_MIXED_BF16 on BF16 MFMA with each FP32
// operand split into two BF16 terms (split-BF16 planes of every finished column of L and of the
// right-hand sides are written as the recursion goes, and of L by columns before the second
// solve), FP32 accumulation; the leaves are FP64 in both.
//
// Cmm carries an exact low-rank term: the reference adds lgs_cst to every entry of each LGS
// xx and yy block (common/kernels/matcov_kernels.cpp:320-321), i.e. Cmm = C0 + c U U^T with U
// the n x p indicator matrix of those 2*nlgs index ranges.  That term makes Cmm
// ill-conditioned (its eigenvalues c*Nsubap sit 4-6 decades above the noise floor; FP32
// Cholesky of Cmm loses definiteness by n ~ 1e4) while C0 stays well conditioned, so:
//   1. C0 (FP32) = Cmm - c U U^T, converted in one pass;
//   2. tall factorization of C0 (FP32 storage) with [Ctm; U^T] riding along, then the second
//      solve: Y = Ctm C0^{-1}, W = U^T C0^{-1}   (low-precision updates, FP64 leaves);
//   3. Woodbury in FP64: S = I/c + W U (p x p, host), R = Y - (Y U) S^{-1} W;
//   4. refinement: E = Ctm - R Cmm (FP64 GEMM on the full Cmm, lower triangle read);
//      stop when ||E||_F <= tol ||Ctm||_F, else D = E C0^{-1} (low-precision solves) corrected
//      by Woodbury, R += D.
// Without a low-rank description (p = 0) the same loop runs on Cmm itself.
// With the environment variable MOAO_DEBUG set, the norms of the intermediate terms and the
// residual of every refinement step are printed to stderr (synchronising after each).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "leaf.hpp"
#include "profile.hpp"
#include "recon.hpp"
#include "stages.hpp"

namespace moao {
namespace {

// C0[r,c] = Cmm[r,c] - c*[r,c in the same low-rank range], lower triangle, to FP32.
__global__ void cmm_to_c0_f32(const double* __restrict__ A, long lda, float* __restrict__ B,
                              long ldb, int n, const int* __restrict__ rid, double cst) {
    const int r = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (r >= n || c >= n || r < c) return;
    double v = A[r + (long)c * lda];
    const int a = rid[r];
    if (a >= 0 && a == rid[c]) v -= cst;
    B[r + (long)c * ldb] = (float)v;
}

__global__ void cvt_f64_f32(const double* __restrict__ a, long lda, float* __restrict__ b,
                            long ldb, int rows, int cols) {
    const int r = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (r < rows && c < cols) b[r + (long)c * ldb] = (float)a[r + (long)c * lda];
}

__global__ void cvt_f32_f64(const float* __restrict__ a, long lda, double* __restrict__ b,
                            long ldb, int rows, int cols) {
    const int r = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (r < rows && c < cols) b[r + (long)c * ldb] = (double)a[r + (long)c * lda];
}

// b += a (FP64)
__global__ void add_f64(const double* __restrict__ a, long lda, double* __restrict__ b, long ldb,
                        int rows, int cols) {
    const int r = blockIdx.x * 64 + (threadIdx.x & 63);
    const int c = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (r < rows && c < cols) b[r + (long)c * ldb] += a[r + (long)c * lda];
}

// Per-workgroup partial sums of squares of a rows x cols matrix (one column strip per group).
__global__ void sumsq_partial(const double* __restrict__ a, long lda, int rows, int cols,
                              double* __restrict__ part) {
    __shared__ double red[256];
    double acc = 0.0;
    const long total = (long)rows * cols;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long c = e / rows, r = e - c * rows;
        const double v = a[r + c * lda];
        acc += v * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

constexpr int kNormBlocks = 1024;

dim3 grid2(int rows, int cols) { return dim3((rows + 63) / 64, (cols + 3) / 4); }

struct Ranges {
    std::vector<long> start, len;
    double cst = 0.0;
    int p() const { return (int)start.size(); }
};

Ranges lowrank_ranges(const moao_tomo_view* t, long n) {
    Ranges R;
    if (!t || t->lgs_cst == 0.0) return R;
    R.cst = t->lgs_cst;
    long off = 0;
    for (int w = 0; w < t->Nw - 1; ++w) {
        const long ns = t->Nsubap[w];
        if (w < t->nlgs) {
            R.start.push_back(2 * off);
            R.len.push_back(ns);
            R.start.push_back(2 * off + ns);
            R.len.push_back(ns);
        }
        off += ns;
    }
    if (2 * off != n) R.start.clear(), R.len.clear();  // view does not describe this Cmm
    return R;
}

// Inverse of the p x p matrix S (column-major) by Gauss-Jordan with partial pivoting; false if
// singular.  S = I/c + W U is used exactly as computed (not symmetrised): the Woodbury identity
// reproduces the inverse of C0_low + c U U^T only when S is formed from the same W that is
// applied, and W from low-precision solves makes W U symmetric only to ~1e-5.
bool small_inverse(const std::vector<double>& S, int p, std::vector<double>& inv) {
    std::vector<double> a(S);
    inv.assign((size_t)p * p, 0.0);
    for (int i = 0; i < p; ++i) inv[i + i * p] = 1.0;
    for (int c = 0; c < p; ++c) {
        int piv = c;
        for (int r = c + 1; r < p; ++r)
            if (std::fabs(a[r + c * p]) > std::fabs(a[piv + c * p])) piv = r;
        if (!(std::fabs(a[piv + c * p]) > 0)) return false;
        if (piv != c)
            for (int k = 0; k < p; ++k) {
                std::swap(a[c + k * p], a[piv + k * p]);
                std::swap(inv[c + k * p], inv[piv + k * p]);
            }
        const double d = a[c + c * p];
        for (int k = 0; k < p; ++k) {
            a[c + k * p] /= d;
            inv[c + k * p] /= d;
        }
        for (int r = 0; r < p; ++r) {
            if (r == c) continue;
            const double f = a[r + c * p];
            if (f == 0.0) continue;
            for (int k = 0; k < p; ++k) {
                a[r + k * p] -= f * a[c + k * p];
                inv[r + k * p] -= f * inv[c + k * p];
            }
        }
    }
    return true;
}

struct MixedWs {
    float* C0;     // n x n
    float* X32;    // (m + p) x n: rows [0,m) right-hand sides, rows [m, m+p) U^T
    float* invL;   // leaves
    double* E;     // m x n residual / correction
    double* Ctm0;  // m x n copy of the right-hand side for the residuals
    double* U;     // n x p indicator
    double* W;     // p x n
    double* Z;     // m x p
    double* Z2;    // m x p
    double* Sinv;  // p x p
    double* WU;    // p x p
    double* part;  // norm partial sums
    int* rid;      // n range ids
    int* info;
};

}  // namespace

// Leading dimension of the split-BF16 planes of an order-n solve.
static long plane_ld(long n) { return (n + 63) / 64 * 64; }

// Byte offsets of the workspace pieces (MixedWs order, then the symmetric diagonal blocks and,
// for the BF16 mode, the planes of L in both orientations and of the right-hand sides) and the
// total size.
static size_t mixed_layout(long n, long m, int p, int mode, std::vector<size_t>* offs) {
    const long xr = m + p;
    const long q = std::max(p, 1);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t off = 0;
    auto take = [&](size_t bytes) {
        if (offs) offs->push_back(off);
        off += al(bytes);
    };
    take(sizeof(float) * n * n);
    take(sizeof(float) * xr * n);
    take(sizeof(float) * recon_invl_doubles(n));
    take(sizeof(double) * m * n);
    take(sizeof(double) * m * n);
    take(sizeof(double) * n * q);
    take(sizeof(double) * q * n);
    take(sizeof(double) * m * q);
    take(sizeof(double) * m * q);
    take(sizeof(double) * q * q);
    take(sizeof(double) * q * q);
    take(sizeof(double) * kNormBlocks);
    take(sizeof(int) * n);
    take(sizeof(int) * 4);
    take(sizeof(double) * gemm_symdiag_elems(n));  // Cmm's symmetric diagonal blocks
    if (mode == MOAO_MIXED_BF16) {
        const long ld = plane_ld(n);
        take(sizeof(unsigned short) * 2 * n * ld);   // L, plane row = row
        take(sizeof(unsigned short) * 2 * n * ld);   // L, plane row = column
        take(sizeof(unsigned short) * 2 * xr * ld);  // right-hand sides
    }
    return off;
}

size_t mixed_workspace_bytes(long n, long m, const moao_tomo_view* lowrank, int mode) {
    return mixed_layout(n, m, lowrank_ranges(lowrank, n).p(), mode, nullptr);
}

int reconstructor_mixed_dev(long n, long m, const double* Cmm, long ldCmm, double* Ctm, long ldCtm,
                            const moao_tomo_view* lowrank, int mode, double tol, int max_iter,
                            double* resid_out, int* iters_out, void* ws, size_t ws_bytes,
                            hipStream_t s) {
    if (n < 1 || m < 1 || ldCmm < n || ldCtm < m || !Cmm || !Ctm || max_iter < 0 ||
        n >= (1L << 31) || m >= (1L << 31) ||
        (mode != MOAO_MIXED_FP32 && mode != MOAO_MIXED_BF16)) {
        set_error("reconstructor_mixed: invalid arguments");
        return MOAO_ERR_ARG;
    }
    const int split = mode == MOAO_MIXED_BF16 ? 1 : 0;
    const bool debug = std::getenv("MOAO_DEBUG") != nullptr;
    const Ranges rg = lowrank_ranges(lowrank, n);
    const int p = rg.p();
    const long xr = m + p;  // rows of the FP32 right-hand-side block
    std::vector<size_t> offs;
    const size_t total = mixed_layout(n, m, p, mode, &offs);
    if (!ws || ws_bytes < total) {
        set_error("reconstructor_mixed: workspace too small");
        return MOAO_ERR_ARG;
    }
    char* b = static_cast<char*>(ws);
    MixedWs w{(float*)(b + offs[0]), (float*)(b + offs[1]), (float*)(b + offs[2]),
              (double*)(b + offs[3]), (double*)(b + offs[4]), (double*)(b + offs[5]),
              (double*)(b + offs[6]), (double*)(b + offs[7]), (double*)(b + offs[8]),
              (double*)(b + offs[9]), (double*)(b + offs[10]), (double*)(b + offs[11]),
              (int*)(b + offs[12]), (int*)(b + offs[13])};
    double* const symd = (double*)(b + offs[14]);
    const int N = (int)n, M = (int)m;
    SplitWs sw{};
    const SplitWs* swp = nullptr;
    if (split) {
        const long ld = plane_ld(n);
        sw = SplitWs{(unsigned short*)(b + offs[15]), (unsigned short*)(b + offs[16]),
                     (unsigned short*)(b + offs[17]), ld, n * ld, n * ld, xr * ld};
        swp = &sw;
    }

    auto norm = [&](const double* a, long lda, int rows, int cols, double* out) -> int {
        hipLaunchKernelGGL(sumsq_partial, dim3(kNormBlocks), dim3(256), 0, s, a, lda, rows, cols,
                           w.part);
        MOAO_HIP_TRY(hipGetLastError());
        std::vector<double> h(kNormBlocks);
        MOAO_HIP_TRY(hipMemcpyAsync(h.data(), w.part, sizeof(double) * kNormBlocks,
                                    hipMemcpyDeviceToHost, s));
        MOAO_HIP_TRY(hipStreamSynchronize(s));
        double acc = 0;
        for (double v : h) acc += v;
        *out = std::sqrt(acc);
        return MOAO_OK;
    };

    // --- range ids, U, U^T
    {
        std::vector<int> rid(n, -1);
        std::vector<double> U((size_t)n * std::max(p, 1), 0.0);
        std::vector<float> Ut((size_t)std::max(p, 1) * n, 0.0f);
        for (int q = 0; q < p; ++q)
            for (long k = rg.start[q]; k < rg.start[q] + rg.len[q]; ++k) {
                rid[k] = q;
                U[k + (size_t)q * n] = 1.0;
            }
        MOAO_HIP_TRY(hipMemcpyAsync(w.rid, rid.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        if (p) {
            MOAO_HIP_TRY(hipMemcpyAsync(w.U, U.data(), sizeof(double) * n * p,
                                        hipMemcpyHostToDevice, s));
            // U^T rows live below the m right-hand-side rows of X32 (leading dimension xr)
            std::vector<float> Xu((size_t)p * n);
            for (int q = 0; q < p; ++q)
                for (long k = 0; k < n; ++k) Xu[q + (size_t)k * p] = (float)U[k + (size_t)q * n];
            MOAO_HIP_TRY(hipMemcpy2DAsync(w.X32 + m, xr * sizeof(float), Xu.data(),
                                          p * sizeof(float), p * sizeof(float), n,
                                          hipMemcpyHostToDevice, s));
        }
        MOAO_HIP_TRY(hipStreamSynchronize(s));  // host vectors go out of scope
    }

    // --- 1. C0 in FP32; right-hand sides Ctm in FP32
    {
        ProfScope ps(kProfConvert, 12.0 * n * (n + 1) / 2, s);
        hipLaunchKernelGGL(cmm_to_c0_f32, grid2(N, N), dim3(256), 0, s, Cmm, ldCmm, w.C0, n, N,
                           w.rid, rg.cst);
        hipLaunchKernelGGL(cvt_f64_f32, grid2(M, N), dim3(256), 0, s, Ctm, ldCtm, w.X32, xr, M, N);
        MOAO_HIP_TRY(hipGetLastError());
    }
    double ctm_norm = 0;
    if (int rc = norm(Ctm, ldCtm, M, N, &ctm_norm)) return rc;
    // keep the original Ctm for the residuals
    MOAO_HIP_TRY(hipMemcpy2DAsync(w.Ctm0, m * sizeof(double), Ctm, ldCtm * sizeof(double),
                                  m * sizeof(double), n, hipMemcpyDeviceToDevice, s));

    // --- 2. FP32 factorization with [Ctm; U^T] riding along, second solve
    std::vector<RowGroup<float>> all{{w.X32, xr, (int)xr}};
    if (int rc = potrf_tall<float>(n, w.C0, n, all, w.invL, w.info, s, swp)) return rc;
    int hinfo = 0;
    MOAO_HIP_TRY(hipMemcpyAsync(&hinfo, w.info, sizeof(int), hipMemcpyDeviceToHost, s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    if (hinfo) {
        set_error("reconstructor_mixed: FP32 factorization of C0 failed");
        return hinfo;
    }
    // the second solve reads L by columns: its planes with plane row = column of L
    if (split)
        if (int rc = split_planes(w.C0, n, N, N, sw.lq, sw.ld, sw.lqs, 0, s)) return rc;
    if (int rc = trsm_right_lower_notrans<float>(n, w.C0, n, w.invL, all, s, nullptr, swp))
        return rc;

    // --- 3. Woodbury terms
    std::vector<double> Sinv;
    if (p) {
        hipLaunchKernelGGL(cvt_f32_f64, grid2(p, N), dim3(256), 0, s, w.X32 + m, xr, w.W, p, p, N);
        MOAO_HIP_TRY(hipGetLastError());
        auto g = gemm_args<double>(p, p, N, 1.0, w.W, p, w.U, n, 0, 0.0, w.WU, p);
        if (int rc = gemm_f64(g, s)) return rc;
        std::vector<double> S((size_t)p * p);
        MOAO_HIP_TRY(hipMemcpyAsync(S.data(), w.WU, sizeof(double) * p * p, hipMemcpyDeviceToHost,
                                    s));
        MOAO_HIP_TRY(hipStreamSynchronize(s));
        for (int q = 0; q < p; ++q) S[q + q * p] += 1.0 / rg.cst;
        if (!small_inverse(S, p, Sinv)) {
            set_error("reconstructor_mixed: Woodbury capacitance matrix is singular");
            return MOAO_ERR_NOCONV;
        }
        if (debug) {
            double tr = 0, trs = 0;
            for (int q = 0; q < p; ++q) tr += S[q + q * p], trs += Sinv[q + q * p];
            std::fprintf(stderr, "[mixed] p=%d trace(S)=%.6e trace(S^-1)=%.6e\n", p, tr, trs);
        }
        MOAO_HIP_TRY(hipMemcpyAsync(w.Sinv, Sinv.data(), sizeof(double) * p * p,
                                    hipMemcpyHostToDevice, s));
    }
    // Y (FP64) -> X; X -= (X U) S^{-1} W
    auto woodbury = [&](double* X, long ldx) -> int {
        if (!p) return MOAO_OK;
        auto g = gemm_args<double>(M, p, N, 1.0, X, ldx, w.U, n, 0, 0.0, w.Z, m);
        if (int rc = gemm_f64(g, s)) return rc;
        g = gemm_args<double>(M, p, p, 1.0, w.Z, m, w.Sinv, p, 0, 0.0, w.Z2, m);
        if (int rc = gemm_f64(g, s)) return rc;
        g = gemm_args<double>(M, N, p, -1.0, w.Z2, m, w.W, p, 0, 1.0, X, ldx);
        return gemm_f64(g, s);
    };
    hipLaunchKernelGGL(cvt_f32_f64, grid2(M, N), dim3(256), 0, s, w.X32, xr, Ctm, ldCtm, M, N);
    MOAO_HIP_TRY(hipGetLastError());
    if (debug) {
        double a = 0, b2 = 0, c2 = 0;
        norm(Ctm, ldCtm, M, N, &a);
        if (p) norm(w.W, p, p, N, &b2);
        if (p) norm(w.U, n, N, p, &c2);
        std::fprintf(stderr, "[mixed] |Ctm|=%.6e |Y|=%.6e |W|=%.6e |U|=%.6e\n", ctm_norm, a, b2, c2);
    }
    if (int rc = woodbury(Ctm, ldCtm)) return rc;  // Ctm now holds R_0
    if (debug && p) {
        double a = 0, b2 = 0, c2 = 0;
        norm(w.Z, m, M, p, &a);
        norm(w.Z2, m, M, p, &b2);
        norm(Ctm, ldCtm, M, N, &c2);
        std::fprintf(stderr, "[mixed] |Z|=%.6e |Z2|=%.6e |R0|=%.6e\n", a, b2, c2);
    }

    // --- 4. refinement (the residual's symmetric operand Cmm feeds the direct-to-LDS GEMM)
    if (int rc = gemm_symdiag(Cmm, ldCmm, N, symd, s)) return rc;
    double rel = 0;
    int it = 0;
    int status = MOAO_OK;
    for (;; ++it) {
        MOAO_HIP_TRY(hipMemcpyAsync(w.E, w.Ctm0, sizeof(double) * m * n, hipMemcpyDeviceToDevice,
                                    s));
        auto g = gemm_args<double>(M, N, N, -1.0, Ctm, ldCtm, Cmm, ldCmm, 2, 1.0, w.E, m);
        g.symdiag = symd;
        if (int rc = gemm_f64(g, s)) { status = rc; break; }
        double en = 0;
        if (int rc = norm(w.E, m, M, N, &en)) { status = rc; break; }
        rel = ctm_norm > 0 ? en / ctm_norm : en;
        if (debug) std::fprintf(stderr, "[mixed] step %d relative residual %.3e\n", it, rel);
        if (rel <= tol) break;
        if (it >= max_iter) { status = MOAO_ERR_NOCONV; break; }
        // D = E C0^{-1} (FP32 solves), Woodbury-corrected in FP64, R += D
        hipLaunchKernelGGL(cvt_f64_f32, grid2(M, N), dim3(256), 0, s, w.E, m, w.X32, xr, M, N);
        std::vector<RowGroup<float>> rhs{{w.X32, xr, M}};
        if (int rc = trsm_right_lower_trans<float>(n, w.C0, n, w.invL, rhs, s, nullptr, swp)) {
            status = rc;
            break;
        }
        if (int rc = trsm_right_lower_notrans<float>(n, w.C0, n, w.invL, rhs, s, nullptr, swp)) {
            status = rc;
            break;
        }
        hipLaunchKernelGGL(cvt_f32_f64, grid2(M, N), dim3(256), 0, s, w.X32, xr, w.E, m, M, N);
        if (int rc = woodbury(w.E, m)) { status = rc; break; }
        hipLaunchKernelGGL(add_f64, grid2(M, N), dim3(256), 0, s, w.E, m, Ctm, ldCtm, M, N);
        MOAO_HIP_TRY(hipGetLastError());
    }
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    if (resid_out) *resid_out = rel;
    if (iters_out) *iters_out = it;
    if (status == MOAO_ERR_NOCONV)
        set_error("reconstructor_mixed: refinement stopped at relative residual " +
                  std::to_string(rel));
    return status;
}

}  // namespace moao

extern "C" {

int moao_hip_reconstructor_mixed_dev(long nmeas, long nmeasts, const double* dCmm, long ldCmm,
                                     double* dCtm, long ldCtm, const moao_tomo_view* lowrank,
                                     int mode, double tol, int max_iter, double* resid,
                                     int* iters, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (mode != MOAO_MIXED_FP32 && mode != MOAO_MIXED_BF16) {
        moao::set_error("reconstructor_mixed: unsupported mode");
        return MOAO_ERR_ARG;
    }
    if (nmeas < 1 || nmeasts < 1 || nmeas >= (1L << 31) || nmeasts >= (1L << 31)) {
        moao::set_error("reconstructor_mixed: invalid arguments");
        return MOAO_ERR_ARG;
    }
    // workspace of this call only, from the stream-ordered pool
    hipStream_t s = moao::pick_stream(stream);
    const size_t bytes = moao::mixed_workspace_bytes(nmeas, nmeasts, lowrank, mode);
    void* ws = nullptr;
    if (hipMallocAsync(&ws, bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        moao::set_error("reconstructor_mixed: workspace allocation of " + std::to_string(bytes) +
                        " B failed");
        return MOAO_ERR_NOMEM;
    }
    int rc = moao::reconstructor_mixed_dev(nmeas, nmeasts, dCmm, ldCmm, dCtm, ldCtm, lowrank, mode,
                                           tol, max_iter, resid, iters, ws, bytes, s);
    (void)hipFreeAsync(ws, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc == MOAO_OK) {
        moao::set_error("reconstructor_mixed: stream synchronisation failed");
        rc = MOAO_ERR_HIP;
    }
    return rc;
}

}  // extern "C"
