// Cee / Cvv error covariances (MOAO_LAPACK::Cee_Cvv, lapack/compute/cee_cvv.cpp:50-124) on gfx950.
//
// The reference chain and what runs here, all FP64 on the MFMA GEMM engine:
//   1 dsyr2k(L,N, -1, Cpm, R, 1, Cpp)       X = Cpm R^T (one GEMM) and Cpp -= X, Cpp -= X^T on
//                                           the lower triangle (one elementwise kernel); two
//                                           lower-trapezoid GEMMs when X has no scratch
//   2 dsymm(R,L, 1, Cmm, R, 0, Cpm)         GEMM with B = Cmm read from its lower triangle
//   3-5 dlaset(U) + dlacpy(L) + in-place geadd('T') + dscal(0.5) on the diagonal
//                                           one kernel: Cee = lower(Cpp) mirrored (the net
//                                           effect of those four calls)
//   6 dgemm(N,T, 1, Cpm, R, 1, Cee)         GEMM
//   7 dsymm(R,L, 1, Cee, Dx, 0, Tmp)        GEMM with B = Cee read from its lower triangle
//   8 dgemm(N,T, 1, Tmp, Dx, 0, Cvv)        GEMM
// so Cee = Ctt - Ctm R^T - R Ctm^T + R Cmm R^T and Cvv = Dx Cee Dx^T, with the reference's side
// effects on Cpp (lower part = step 1 result) and Cpm (= R Cmm).  Unlike the reference, every
// leading-dimension argument is honoured (the reference passes nmeasts/nmeas/nact regardless).
//
// The batched entry runs the chain for ngal targets (the observation step's per-target calls)
// with each step's products of up to kMaxSeg targets in one launch (GEMM segments with their own
// A, C and, for steps 1 and 6, B; Cmm and Dx shared): at m = 2,336 one target's products leave
// most of a launch's last wave of tiles empty, eight targets' fill it.  Each element goes through
// the same kernel arithmetic as in a call for its target alone, so the outputs are identical.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "profile.hpp"

namespace moao {
namespace {

// Cee[i, j] = Cpp[max(i,j), min(i,j)]  (laset + lacpy + geadd('T') + scal on the diagonal)
__global__ void mirror_lower(const double* __restrict__ src, long lds, double* __restrict__ dst,
                             long ldd, int m) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= m || j >= m) return;
    dst[i + (long)j * ldd] = i >= j ? src[i + (long)j * lds] : src[j + (long)i * lds];
}

// Cpp[i, j] = (Cpp[i, j] - X[i, j]) - X[j, i] for i >= j, target blockIdx.z: the lower part of
// Cpp - X - X^T, rounded in the order of two accumulating products Cpp -= X, Cpp -= X^T.
__global__ void syr2_apply(const double* __restrict__ X, long xstride, double* Cpp, long ldCpp,
                           long cstride, int m) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= m || j >= m || i < j) return;
    const double* x = X + blockIdx.z * xstride;
    double* c = Cpp + blockIdx.z * cstride + i + (long)j * ldCpp;
    const double v = *c - x[i + (long)j * m];
    *c = v - x[j + (long)i * m];
}

}  // namespace

int cee_cvv_batched_dev(long n, long m, long nact, long ngal, const double* Cmm, long ldCmm,
                        double* Cpp, long ldCpp, double* Cpm, long ldCpm, const double* R,
                        long ldR, const double* Dx, long ldDx, double* Cee, long ldCee,
                        double* Cvv, long ldCvv, double* Tmp, long ldTmp, hipStream_t s) {
    if (n < 1 || m < 1 || nact < 0 || ngal < 1 || ldCmm < n || ldCpp < m || ldCpm < m ||
        ldR < m || (nact > 0 && ldDx < nact) || ldCee < m ||
        (nact > 0 && (ldCvv < nact || ldTmp < nact)) || n >= (1L << 31) || m >= (1L << 31) ||
        nact >= (1L << 31)) {
        set_error("cee_cvv: invalid arguments");
        return MOAO_ERR_ARG;
    }
    const int N = (int)n, M = (int)m, A = (int)nact;
    // target g's blocks: one block stride (leading dimension x columns) after target g-1's
    auto Cpp_ = [&](long g) { return Cpp + g * ldCpp * m; };
    auto Cpm_ = [&](long g) { return Cpm + g * ldCpm * n; };
    auto R_ = [&](long g) { return R + g * ldR * n; };
    auto Cee_ = [&](long g) { return Cee + g * ldCee * m; };
    auto Cvv_ = [&](long g) { return Cvv + g * ldCvv * nact; };
    auto Tmp_ = [&](long g) { return Tmp + g * ldTmp * m; };
    // Stream-ordered scratch for the symmetric operands' diagonal blocks (steps 2 and 7), so
    // calls on different streams never share it; without it those steps use the
    // register-staged kernel.
    struct Scratch {
        double* p = nullptr;
        hipStream_t s;
        ~Scratch() { if (p) (void)hipFreeAsync(p, s); }
    } sd{nullptr, s}, sx{nullptr, s};
    if (hipMallocAsync(reinterpret_cast<void**>(&sd.p),
                       sizeof(double) * gemm_symdiag_elems(std::max(n, m)), s) != hipSuccess) {
        sd.p = nullptr;
        (void)hipGetLastError();
    }
    // step 1's product X = Cpm R^T for up to kMaxSeg targets (m x m each); without it step 1
    // runs as two lower-trapezoid products
    const long nx = std::min<long>(ngal, kMaxSeg);
    if (hipMallocAsync(reinterpret_cast<void**>(&sx.p), sizeof(double) * m * m * nx, s) !=
        hipSuccess) {
        sx.p = nullptr;
        (void)hipGetLastError();
    }
    // One launch per step for up to kMaxSeg targets: segment i is target g0 + i, with its own
    // A, C and (steps 1 and 6) B.
    using Seg = GemmSeg<double>;
    auto batch = [&](GemmArgs<double> g, long g0, int ns, auto&& seg) {
        g.nseg = ns;
        for (int i = 0; i < ns; ++i) g.seg[i] = seg(g0 + i);
        return gemm_f64(g, s);
    };
    for (long g0 = 0; g0 < ngal; g0 += kMaxSeg) {
        const int ns = (int)std::min<long>(kMaxSeg, ngal - g0);
        // 1: Cpp(lower) -= Cpm R^T + R Cpm^T.  (R Cpm^T)[i, j] = X[j, i] bit for bit with
        // X = Cpm R^T (the same products, accumulated in the same k order), so one full
        // product and an elementwise pass give exactly the two lower-trapezoid products'
        // result, with all of one launch's tiles busy instead of half of two launches'.
        if (sx.p) {
            auto g = gemm_args<double>(M, M, N, 1.0, nullptr, ldCpm, R_(g0), ldR, 1, 0.0, nullptr,
                                       m);
            if (int rc = batch(g, g0, ns, [&](long t) {
                    Seg q{Cpm_(t), sx.p + (t - g0) * m * m, ldCpm, m, M};
                    q.B = R_(t);
                    return q;
                }))
                return rc;
            ProfScope ps(kProfOther, 24.0 * m * m / 2 * ns, s);
            dim3 grid((M + 63) / 64, (M + 3) / 4, ns);
            hipLaunchKernelGGL(syr2_apply, grid, dim3(256), 0, s, sx.p, m * m, Cpp_(g0),
                               ldCpp, ldCpp * m, M);
            MOAO_HIP_TRY(hipGetLastError());
            continue;
        }
        auto g = gemm_args<double>(M, M, N, -1.0, nullptr, ldCpm, R_(g0), ldR, 1, 1.0, nullptr,
                                   ldCpp);
        g.tri = 1;
        g.tri_all = 1;
        if (int rc = batch(g, g0, ns, [&](long t) {
                Seg q{Cpm_(t), Cpp_(t), ldCpm, ldCpp, M};
                q.B = R_(t);
                return q;
            }))
            return rc;
        g = gemm_args<double>(M, M, N, -1.0, nullptr, ldR, Cpm_(g0), ldCpm, 1, 1.0, nullptr,
                              ldCpp);
        g.tri = 1;
        g.tri_all = 1;
        if (int rc = batch(g, g0, ns, [&](long t) {
                Seg q{R_(t), Cpp_(t), ldR, ldCpp, M};
                q.B = Cpm_(t);
                return q;
            }))
            return rc;
    }
    // 2: Cpm = R sym(Cmm), every target against the one Cmm
    if (sd.p)
        if (int rc = gemm_symdiag(Cmm, ldCmm, N, sd.p, s)) return rc;
    for (long g0 = 0; g0 < ngal; g0 += kMaxSeg) {
        const int ns = (int)std::min<long>(kMaxSeg, ngal - g0);
        auto g = gemm_args<double>(M, N, N, 1.0, nullptr, ldR, Cmm, ldCmm, 2, 0.0, nullptr, ldCpm);
        g.symdiag = sd.p;
        if (int rc = batch(g, g0, ns, [&](long t) { return Seg{R_(t), Cpm_(t), ldR, ldCpm, M}; }))
            return rc;
    }
    for (long g0 = 0; g0 < ngal; g0 += kMaxSeg) {
        const int ns = (int)std::min<long>(kMaxSeg, ngal - g0);
        // 3-5: Cee = lower(Cpp) mirrored
        {
            ProfScope ps(kProfOther, 16.0 * m * m * ns, s);
            dim3 grid((M + 63) / 64, (M + 3) / 4);
            for (int i = 0; i < ns; ++i)
                hipLaunchKernelGGL(mirror_lower, grid, dim3(256), 0, s, Cpp_(g0 + i), ldCpp,
                                   Cee_(g0 + i), ldCee, M);
            MOAO_HIP_TRY(hipGetLastError());
        }
        // 6: Cee += Cpm R^T
        auto g = gemm_args<double>(M, M, N, 1.0, nullptr, ldCpm, R_(g0), ldR, 1, 1.0, nullptr,
                                   ldCee);
        if (int rc = batch(g, g0, ns, [&](long t) {
                Seg q{Cpm_(t), Cee_(t), ldCpm, ldCee, M};
                q.B = R_(t);
                return q;
            }))
            return rc;
    }
    if (A == 0) return MOAO_OK;
    // 7: Tmp = Dx sym(lower(Cee)), one launch per target (each has its own symmetric operand)
    for (long t = 0; t < ngal; ++t) {
        auto g = gemm_args<double>(A, M, M, 1.0, Dx, ldDx, Cee_(t), ldCee, 2, 0.0, Tmp_(t), ldTmp);
        if (sd.p) {
            if (int rc = gemm_symdiag(Cee_(t), ldCee, M, sd.p, s)) return rc;
            g.symdiag = sd.p;
        }
        if (int rc = gemm_f64(g, s)) return rc;
    }
    // 8: Cvv = Tmp Dx^T, every target against the one Dx
    for (long g0 = 0; g0 < ngal; g0 += kMaxSeg) {
        const int ns = (int)std::min<long>(kMaxSeg, ngal - g0);
        auto g = gemm_args<double>(A, A, M, 1.0, nullptr, ldTmp, Dx, ldDx, 1, 0.0, nullptr, ldCvv);
        if (int rc = batch(g, g0, ns, [&](long t) { return Seg{Tmp_(t), Cvv_(t), ldTmp, ldCvv, A}; }))
            return rc;
    }
    return MOAO_OK;
}

int cee_cvv_dev(long n, long m, long nact, const double* Cmm, long ldCmm, double* Cpp,
                long ldCpp, double* Cpm, long ldCpm, const double* R, long ldR,
                const double* Dx, long ldDx, double* Cee, long ldCee, double* Cvv, long ldCvv,
                double* Tmp, long ldTmp, hipStream_t s) {
    return cee_cvv_batched_dev(n, m, nact, 1, Cmm, ldCmm, Cpp, ldCpp, Cpm, ldCpm, R, ldR, Dx,
                               ldDx, Cee, ldCee, Cvv, ldCvv, Tmp, ldTmp, s);
}

}  // namespace moao

extern "C" {

int moao_hip_cee_cvv_batched_dev(long nmeas, long nmeasts, long nact, long ngal,
                                 const double* dCmm, long ldCmm, double* dCpp, long ldCpp,
                                 double* dCpm, long ldCpm, const double* dR, long ldR,
                                 const double* dDx, long ldDx, double* dCee, long ldCee,
                                 double* dCvv, long ldCvv, double* dTmp, long ldTmp,
                                 void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    return moao::cee_cvv_batched_dev(nmeas, nmeasts, nact, ngal, dCmm, ldCmm, dCpp, ldCpp, dCpm,
                                     ldCpm, dR, ldR, dDx, ldDx, dCee, ldCee, dCvv, ldCvv, dTmp,
                                     ldTmp, moao::pick_stream(stream));
}

int moao_hip_cee_cvv_dev(long nmeas, long nmeasts, long nact, const double* dCmm, long ldCmm,
                         double* dCpp, long ldCpp, double* dCpm, long ldCpm, const double* dR,
                         long ldR, const double* dDx, long ldDx, double* dCee, long ldCee,
                         double* dCvv, long ldCvv, double* dTmp, long ldTmp, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    return moao::cee_cvv_dev(nmeas, nmeasts, nact, dCmm, ldCmm, dCpp, ldCpp, dCpm, ldCpm, dR,
                             ldR, dDx, ldDx, dCee, ldCee, dCvv, ldCvv, dTmp, ldTmp,
                             moao::pick_stream(stream));
}

// Host-buffer drop-in: stages every operand through device memory and copies back Cpp, Cpm,
// Cee, Cvv and Tmp (all of which the reference writes).
int moao_hip_cee_cvv(long nmeas, long nmeasts, long nact, double* Cmm, long ldCmm, double* Cpp,
                     long ldCpp, double* Cpm, long ldCpm, double* R, long ldR, double* Dx,
                     long ldDx, double* Cee, long ldCee, double* Cvv, long ldCvv, double* Tmp,
                     long ldTmp) {
    if (int rc = moao::ensure_init()) return rc;
    if (!Cmm || !Cpp || !Cpm || !R || !Cee || (nact > 0 && (!Dx || !Cvv || !Tmp)) ||
        nmeas < 1 || nmeasts < 1 || nact < 0) {
        moao::set_error("cee_cvv: invalid arguments");
        return MOAO_ERR_ARG;
    }
    struct Op { double* h; long ld, rows, cols; bool in, out; size_t off; };
    Op ops[] = {{Cmm, ldCmm, nmeas, nmeas, true, false, 0},
                {Cpp, ldCpp, nmeasts, nmeasts, true, true, 0},
                {Cpm, ldCpm, nmeasts, nmeas, true, true, 0},
                {R, ldR, nmeasts, nmeas, true, false, 0},
                {Dx, ldDx, nact, nmeasts, nact > 0, false, 0},
                {Cee, ldCee, nmeasts, nmeasts, false, true, 0},
                {Cvv, ldCvv, nact, nact, false, nact > 0, 0},
                {Tmp, ldTmp, nact, nmeasts, false, nact > 0, 0}};
    size_t total = 0;
    for (Op& o : ops) {
        o.off = total;
        const size_t span = o.rows > 0 && o.cols > 0 ? (size_t)(o.cols - 1) * o.ld + o.rows : 0;
        total += (span * sizeof(double) + 255) & ~size_t(255);
    }
    moao::Context& c = moao::ctx();
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = c.host_stage.ensure(total + 256)) return rc;
    char* base = c.host_stage.as<char>();
    hipStream_t s = c.stream;
    auto span = [](const Op& o) {
        return o.rows > 0 && o.cols > 0 ? ((size_t)(o.cols - 1) * o.ld + o.rows) * sizeof(double)
                                         : 0;
    };
    // Outputs that leave gaps between columns (ld > rows) start from the caller's contents.
    for (Op& o : ops)
        if ((o.in || (o.out && o.ld != o.rows)) && span(o))
            MOAO_HIP_TRY(hipMemcpyAsync(base + o.off, o.h, span(o), hipMemcpyHostToDevice, s));
    auto D = [&](int i) { return (double*)(base + ops[i].off); };
    int rc = moao::cee_cvv_dev(nmeas, nmeasts, nact, D(0), ldCmm, D(1), ldCpp, D(2), ldCpm, D(3),
                               ldR, D(4), ldDx, D(5), ldCee, D(6), ldCvv, D(7), ldTmp, s);
    if (rc) return rc;
    for (Op& o : ops)
        if (o.out && span(o))
            MOAO_HIP_TRY(hipMemcpyAsync(o.h, base + o.off, span(o), hipMemcpyDeviceToHost, s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return MOAO_OK;
}

}  // extern "C"
