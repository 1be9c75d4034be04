// Reconstructor R = Ctm * Cmm^{-1} (lapack/compute/reconstructor.cpp:32-58) on gfx950.
//
// The reference calls dpotrf('L', Cmm), then per target dtrsm(Right, Lower, Trans) and
// dtrsm(Right, Lower, NoTrans) on Ctm.  Here:
//  * Cholesky and the first solve are one recursive factorization of the "tall" matrix
//    [Cmm; Ctm_0; ...; Ctm_{ngal-1}] (only the top n x n block is SPD): every update GEMM also
//    updates the Ctm rows, so X = Ctm * L^{-T} falls out of the factorization;
//  * the second solve Y * L = X runs backwards by the same recursion;
//  * leaves (kLeaf columns) are factored and inverted in LDS; their panel solves are in-place
//    GEMMs with the inverse, so every flop except the leaves runs on MFMA.
// The split point of every recursion is a multiple of kLeaf, so the top-level updates are the
// largest GEMMs (K = n/2) and the deep levels stay cheap.
#include <hip/hip_runtime.h>

#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "leaf.hpp"
#include "recon.hpp"

namespace moao {

size_t recon_invl_doubles(long n) {
    const long nl = (n + kLeaf - 1) / kLeaf;
    return (size_t)nl * kLeaf * kLeaf;
}

namespace {

struct Problem {
    long n, m, ngal;
    double* Cmm;
    long ldA;
    double* Ctm;
    long ldB;
    double* invL;
    int* info;
    hipStream_t s;
    double* ctm(long g) const { return Ctm + g * n * m; }
    double* cmm(long r, long c) const { return Cmm + r + c * ldA; }
};

long split(long nn) {
    long n1 = ((nn / 2 + kLeaf - 1) / kLeaf) * kLeaf;
    if (n1 >= nn) n1 = nn - kLeaf;
    return n1;
}

// Target groups of one launch: segment 0 is the Cmm rows (if any), then Ctm targets.
template <typename F>
int for_target_chunks(const Problem& p, int first_free, F&& fn) {
    long g = 0;
    do {
        const int cap = kMaxSeg - first_free;
        const long cnt = std::min<long>(cap, p.ngal - g);
        if (int rc = fn(g, (int)cnt)) return rc;
        g += cnt;
        first_free = 0;
    } while (g < p.ngal);
    return MOAO_OK;
}

// Factor columns [j0, j0+nn) of the tall matrix (all earlier columns already applied).
int potrf_rec(const Problem& p, long j0, long nn) {
    if (nn <= kLeaf) {
        double* inv = p.invL + (j0 / kLeaf) * (long)kLeaf * kLeaf;
        if (int rc = leaf_potrf_inv_launch(p.cmm(j0, j0), p.ldA, (int)nn, inv, p.info, (int)j0, p.s))
            return rc;
        // Panel: rows below the leaf (Cmm) and every Ctm row, times inv(L)^T, in place.
        const long below = p.n - j0 - nn;
        bool cmm_done = below <= 0;
        return for_target_chunks(p, cmm_done ? 0 : 1, [&](long g0, int cnt) {
            GemmArgs<double> a{};
            int ns = 0;
            if (!cmm_done) {
                a.seg[ns++] = {p.cmm(j0 + nn, j0), p.cmm(j0 + nn, j0), p.ldA, p.ldA, (int)below};
                cmm_done = true;
            }
            for (int q = 0; q < cnt; ++q) {
                double* c = p.ctm(g0 + q) + j0 * p.ldB;
                a.seg[ns++] = {c, c, p.ldB, p.ldB, (int)p.m};
            }
            a.nseg = ns;
            a.B = inv;
            a.ldb = kLeaf;
            a.N = (int)nn;
            a.K = (int)nn;
            a.transb = 1;
            a.alpha = 1.0;
            a.beta = 0.0;
            return gemm_f64(a, p.s);
        });
    }
    const long n1 = split(nn);
    if (int rc = potrf_rec(p, j0, n1)) return rc;
    // Columns [j0+n1, j0+nn) of all rows below j0+n1 -= rows * Cmm[j0+n1:j0+nn, j0:j0+n1]^T
    const long r0 = j0 + n1;
    bool cmm_done = false;
    int rc = for_target_chunks(p, 1, [&](long g0, int cnt) {
        GemmArgs<double> a{};
        int ns = 0;
        if (!cmm_done) {
            a.seg[ns++] = {p.cmm(r0, j0), p.cmm(r0, r0), p.ldA, p.ldA, (int)(p.n - r0)};
            a.tri = 1;
            a.tri_off = 0;
            cmm_done = true;
        }
        for (int q = 0; q < cnt; ++q) {
            double* c = p.ctm(g0 + q);
            a.seg[ns++] = {c + j0 * p.ldB, c + r0 * p.ldB, p.ldB, p.ldB, (int)p.m};
        }
        a.nseg = ns;
        a.B = p.cmm(r0, j0);
        a.ldb = p.ldA;
        a.N = (int)(nn - n1);
        a.K = (int)n1;
        a.transb = 1;
        a.alpha = -1.0;
        a.beta = 1.0;
        return gemm_f64(a, p.s);
    });
    if (rc) return rc;
    return potrf_rec(p, r0, nn - n1);
}

// Solve Y * L = X for columns [j0, j0+nn) given contributions of Y[:, j0+nn:] already removed.
int trsm2_rec(const Problem& p, long j0, long nn) {
    if (nn <= kLeaf) {
        const double* inv = p.invL + (j0 / kLeaf) * (long)kLeaf * kLeaf;
        return for_target_chunks(p, 0, [&](long g0, int cnt) {
            GemmArgs<double> a{};
            for (int q = 0; q < cnt; ++q) {
                double* c = p.ctm(g0 + q) + j0 * p.ldB;
                a.seg[q] = {c, c, p.ldB, p.ldB, (int)p.m};
            }
            a.nseg = cnt;
            a.B = inv;
            a.ldb = kLeaf;
            a.N = (int)nn;
            a.K = (int)nn;
            a.transb = 0;
            a.alpha = 1.0;
            a.beta = 0.0;
            return gemm_f64(a, p.s);
        });
    }
    const long n1 = split(nn);
    if (int rc = trsm2_rec(p, j0 + n1, nn - n1)) return rc;
    // X[:, j0:j0+n1] -= Y[:, j0+n1:j0+nn] * L[j0+n1:j0+nn, j0:j0+n1]
    int rc = for_target_chunks(p, 0, [&](long g0, int cnt) {
        GemmArgs<double> a{};
        for (int q = 0; q < cnt; ++q) {
            double* c = p.ctm(g0 + q);
            a.seg[q] = {c + (j0 + n1) * p.ldB, c + j0 * p.ldB, p.ldB, p.ldB, (int)p.m};
        }
        a.nseg = cnt;
        a.B = p.cmm(j0 + n1, j0);
        a.ldb = p.ldA;
        a.N = (int)n1;
        a.K = (int)(nn - n1);
        a.transb = 0;
        a.alpha = -1.0;
        a.beta = 1.0;
        return gemm_f64(a, p.s);
    });
    if (rc) return rc;
    return trsm2_rec(p, j0, n1);
}

}  // namespace

int reconstructor_f64_dev(long n, long m, double* dCmm, long ldCmm, double* dCtm, long ldCtm,
                          long ngal, double* dinvL, int* dinfo, hipStream_t stream) {
    if (n < 1 || m < 0 || ngal < 1 || ldCmm < n || (m > 0 && ldCtm < m) || !dCmm || !dinvL ||
        !dinfo || n >= (1L << 31) || m >= (1L << 31)) {
        set_error("reconstructor: invalid arguments");
        return MOAO_ERR_ARG;
    }
    Problem p{n, m, ngal, dCmm, ldCmm, dCtm, ldCtm, dinvL, dinfo, stream};
    if (m == 0) p.ngal = 0;
    MOAO_HIP_TRY(hipMemsetAsync(dinfo, 0, sizeof(int), stream));
    if (int rc = potrf_rec(p, 0, n)) return rc;
    int hinfo = 0;
    MOAO_HIP_TRY(hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, stream));
    MOAO_HIP_TRY(hipStreamSynchronize(stream));
    if (hinfo != 0) return hinfo;
    if (m > 0)
        if (int rc = trsm2_rec(p, 0, n)) return rc;
    return MOAO_OK;
}

}  // namespace moao

namespace {

// Device workspace of one FP64 solve: the inverse leaves, then the status word.
int recon_workspace(long n, double** invL, int** info) {
    moao::Context& c = moao::ctx();
    const size_t bytes = moao::recon_invl_doubles(n) * sizeof(double) + 256;
    if (int rc = c.recon_ws.ensure(bytes)) return rc;
    *invL = c.recon_ws.as<double>();
    *info = reinterpret_cast<int*>(c.recon_ws.as<char>() + bytes - 256);
    return MOAO_OK;
}

}  // namespace

extern "C" {

int moao_hip_reconstructor_dev(long nmeas, long nmeasts, double* dCmm, long ldCmm, double* dCtm,
                               long ldCtm, long ngal, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    double* invL;
    int* info;
    std::lock_guard<std::mutex> g(moao::ctx().mu);
    if (int rc = recon_workspace(nmeas, &invL, &info)) return rc;
    return moao::reconstructor_f64_dev(nmeas, nmeasts, dCmm, ldCmm, dCtm, ldCtm, ngal, invL,
                                       info, moao::pick_stream(stream));
}

int moao_hip_reconstructor(long nmeas, long nmeasts, double* Cmm, long ldCmm, double* Ctm,
                           long ldCtm, long ngal) {
    if (int rc = moao::ensure_init()) return rc;
    if (!Cmm || (nmeasts > 0 && !Ctm) || nmeas < 1 || nmeasts < 0 || ngal < 1 ||
        ldCmm < nmeas || (nmeasts > 0 && ldCtm < nmeasts)) {
        moao::set_error("reconstructor: invalid arguments");
        return MOAO_ERR_ARG;
    }
    moao::Context& c = moao::ctx();
    hipStream_t s = c.stream;
    std::lock_guard<std::mutex> g(c.mu);
    const size_t a_bytes = (size_t)ldCmm * nmeas * sizeof(double);
    // Ctm of target g starts at g*nmeas*nmeasts (lapack/compute/reconstructor.cpp:43).
    const size_t b_elems = nmeasts > 0 ? (size_t)(ngal - 1) * nmeas * nmeasts +
                                             (size_t)(nmeas - 1) * ldCtm + nmeasts
                                       : 0;
    if (int rc = c.host_stage.ensure(a_bytes)) return rc;
    if (int rc = c.host_stage2.ensure(b_elems * sizeof(double) + 8)) return rc;
    double* dA = c.host_stage.as<double>();
    double* dB = c.host_stage2.as<double>();
    MOAO_HIP_TRY(hipMemcpyAsync(dA, Cmm, a_bytes, hipMemcpyHostToDevice, s));
    if (b_elems)
        MOAO_HIP_TRY(hipMemcpyAsync(dB, Ctm, b_elems * sizeof(double), hipMemcpyHostToDevice, s));
    double* invL;
    int* info;
    if (int rc = recon_workspace(nmeas, &invL, &info)) return rc;
    int rc = moao::reconstructor_f64_dev(nmeas, nmeasts, dA, ldCmm, dB, ldCtm, ngal, invL, info, s);
    if (rc < 0) return rc;
    // dpotrf overwrites Cmm with its factor even when it stops early; R only on success.
    MOAO_HIP_TRY(hipMemcpyAsync(Cmm, dA, a_bytes, hipMemcpyDeviceToHost, s));
    if (rc == 0 && b_elems)
        MOAO_HIP_TRY(hipMemcpyAsync(Ctm, dB, b_elems * sizeof(double), hipMemcpyDeviceToHost, s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return rc;
}

}  // extern "C"
