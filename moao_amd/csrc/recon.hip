// Reconstructor R = Ctm * Cmm^{-1} (lapack/compute/reconstructor.cpp:32-58) on gfx950.
//
// The reference calls dpotrf('L', Cmm), then per target dtrsm(Right, Lower, Trans) and
// dtrsm(Right, Lower, NoTrans) on Ctm.  Here:
//  * Cholesky and the first solve are one recursive factorization of the "tall" matrix
//    [Cmm; Ctm_0; ...; Ctm_{ngal-1}] (only the top n x n block is SPD): every update GEMM also
//    updates the Ctm rows, so X = Ctm * L^{-T} falls out of the factorization;
//  * the second solve Y * L = X runs backwards by the same recursion;
//  * leaves (kLeaf columns) are factored and inverted in LDS in FP64; their panel solves are
//    in-place GEMMs with the inverse, so every flop except the leaves runs on MFMA.
// The split point of every recursion is a multiple of kLeaf, so the top-level updates are the
// largest GEMMs (K = n/2) and the deep levels stay cheap.  The same recursion runs on FP32
// storage (FP32 MFMA updates, FP64 leaves) for the mixed-precision path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <mutex>
#include <string>
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

long split(long nn) {
    long n1 = ((nn / 2 + kLeaf - 1) / kLeaf) * kLeaf;
    if (n1 >= nn) n1 = nn - kLeaf;
    return n1;
}

template <typename T>
struct Tall {
    long n;
    T* A;
    long lda;
    const std::vector<RowGroup<T>>* groups;
    T* invL;
    int* info;
    const int* skip;  // device status word: launches are no-ops once it is non-zero
    hipStream_t s;
    const SplitWs* sw;  // FP32 storage: update products on the split-BF16 planes kernel
    T* a(long r, long c) const { return A + r + c * lda; }
    T* inv(long j0) const { return invL + (j0 / kLeaf) * (long)kLeaf * kLeaf; }
};

// Launch one GEMM over [optional lead segment] + every row group, kMaxSeg segments at a time.
// make(q) fills the segment of group q (column offsets applied by the caller).
template <typename T, typename Lead, typename Make>
int over_groups(const std::vector<RowGroup<T>>& groups, GemmArgs<T> proto, bool has_lead,
                Lead&& lead, Make&& make, hipStream_t s) {
    size_t q = 0;
    bool lead_pending = has_lead;
    while (lead_pending || q < groups.size()) {
        GemmArgs<T> a = proto;
        int ns = 0;
        a.tri = 0;
        if (lead_pending) {
            a.seg[ns++] = lead();
            a.tri = proto.tri;
            lead_pending = false;
        }
        while (ns < kMaxSeg && q < groups.size()) {
            if (groups[q].rows > 0) a.seg[ns++] = make(groups[q]);
            ++q;
        }
        a.nseg = ns;
        if (ns)
            if (int rc = gemm(a, s)) return rc;
    }
    return MOAO_OK;
}

// Planes of the columns [j0, j0 + nn) just finalised, for the split-BF16 updates: rows
// [r0, n) of the SPD block (lead) and every row of the (single) row group.
template <typename T>
int split_cols_done(const Tall<T>& p, long r0, long j0, long nn, bool lead) {
    if constexpr (std::is_same<T, float>::value) {
        const SplitWs& w = *p.sw;
        if (lead && p.n > r0)
            if (int rc = split_planes(p.a(r0, j0), p.lda, (int)(p.n - r0), (int)nn,
                                      w.lp + r0 * w.ld + j0, w.ld, w.lps, 1, p.s))
                return rc;
        for (const auto& g : *p.groups)
            if (g.rows > 0)
                if (int rc = split_planes(g.p + j0 * g.ld, g.ld, g.rows, (int)nn, w.xp + j0, w.ld,
                                          w.xps, 1, p.s))
                    return rc;
        return MOAO_OK;
    } else {
        set_error("split-BF16 planes need FP32 storage");
        return MOAO_ERR_ARG;
    }
}

// Factor columns [j0, j0+nn) of the tall matrix (all earlier columns already applied).
template <typename T>
int potrf_rec(const Tall<T>& p, long j0, long nn) {
    const auto& G = *p.groups;
    if (nn <= kLeaf) {
        if (int rc = leaf_potrf_inv_launch<T>(p.a(j0, j0), p.lda, (int)nn, p.inv(j0), p.info,
                                              (int)j0, p.s))
            return rc;
        // Panel: rows below the leaf and every group row, times inv(L)^T, in place.
        const long below = p.n - j0 - nn;
        GemmArgs<T> proto{};
        proto.skip_if = p.skip;
        proto.B = p.inv(j0);
        proto.ldb = kLeaf;
        proto.N = (int)nn;
        proto.K = (int)nn;
        proto.transb = 1;
        proto.alpha = T(1);
        proto.beta = T(0);
        if (int rc = over_groups<T>(
                G, proto, below > 0,
                [&] {
                    return GemmSeg<T>{p.a(j0 + nn, j0), p.a(j0 + nn, j0), p.lda, p.lda,
                                      (int)below};
                },
                [&](const RowGroup<T>& g) {
                    T* c = g.p + j0 * g.ld;
                    return GemmSeg<T>{c, c, g.ld, g.ld, g.rows};
                },
                p.s))
            return rc;
        return p.sw ? split_cols_done(p, j0, j0, nn, true) : MOAO_OK;
    }
    const long n1 = split(nn);
    if (int rc = potrf_rec(p, j0, n1)) return rc;
    // Columns [j0+n1, j0+nn) of every row below j0+n1 -= rows * A[j0+n1:j0+nn, j0:j0+n1]^T
    const long r0 = j0 + n1;
    GemmArgs<T> proto{};
    proto.skip_if = p.skip;
    proto.B = p.a(r0, j0);
    proto.ldb = p.lda;
    proto.N = (int)(nn - n1);
    proto.K = (int)n1;
    proto.transb = 1;
    proto.tri = 1;
    proto.tri_off = 0;
    proto.alpha = T(-1);
    proto.beta = T(1);
    if (p.sw) {
        proto.split = 2;
        proto.bp = p.sw->l_rows(r0, j0);
    }
    int rc = over_groups<T>(
        G, proto, true,
        [&] {
            GemmSeg<T> sg{p.a(r0, j0), p.a(r0, r0), p.lda, p.lda, (int)(p.n - r0)};
            if (p.sw) sg.ap = p.sw->l_rows(r0, j0);
            return sg;
        },
        [&](const RowGroup<T>& g) {
            GemmSeg<T> sg{g.p + j0 * g.ld, g.p + r0 * g.ld, g.ld, g.ld, g.rows};
            if (p.sw) sg.ap = p.sw->x_rows(j0);
            return sg;
        },
        p.s);
    if (rc) return rc;
    return potrf_rec(p, r0, nn - n1);
}

// Y * L = X for columns [j0, j0+nn), contributions of Y[:, j0+nn:] already removed.
template <typename T>
int trsmN_rec(const Tall<T>& p, long j0, long nn) {
    const auto& G = *p.groups;
    auto none = [] { return GemmSeg<T>{}; };
    if (nn <= kLeaf) {
        GemmArgs<T> proto{};
        proto.skip_if = p.skip;
        proto.B = p.inv(j0);
        proto.ldb = kLeaf;
        proto.N = (int)nn;
        proto.K = (int)nn;
        proto.transb = 0;
        proto.alpha = T(1);
        proto.beta = T(0);
        if (int rc = over_groups<T>(G, proto, false, none,
                                    [&](const RowGroup<T>& g) {
                                        T* c = g.p + j0 * g.ld;
                                        return GemmSeg<T>{c, c, g.ld, g.ld, g.rows};
                                    },
                                    p.s))
            return rc;
        return p.sw ? split_cols_done(p, 0, j0, nn, false) : MOAO_OK;
    }
    const long n1 = split(nn);
    if (int rc = trsmN_rec(p, j0 + n1, nn - n1)) return rc;
    // X[:, j0:j0+n1] -= Y[:, j0+n1:j0+nn] * L[j0+n1:j0+nn, j0:j0+n1]
    GemmArgs<T> proto{};
    proto.skip_if = p.skip;
    proto.B = p.a(j0 + n1, j0);
    proto.ldb = p.lda;
    proto.N = (int)n1;
    proto.K = (int)(nn - n1);
    proto.transb = 0;
    proto.alpha = T(-1);
    proto.beta = T(1);
    if (p.sw) {
        proto.split = 2;
        proto.bp = p.sw->l_cols(j0, j0 + n1);
    }
    int rc = over_groups<T>(G, proto, false, none,
                            [&](const RowGroup<T>& g) {
                                GemmSeg<T> sg{g.p + (j0 + n1) * g.ld, g.p + j0 * g.ld, g.ld, g.ld,
                                              g.rows};
                                if (p.sw) sg.ap = p.sw->x_rows(j0 + n1);
                                return sg;
                            },
                            p.s);
    if (rc) return rc;
    return trsmN_rec(p, j0, n1);
}

// X * L^T = B for columns [j0, j0+nn), contributions of X[:, :j0] already removed.
template <typename T>
int trsmT_rec(const Tall<T>& p, long j0, long nn) {
    const auto& G = *p.groups;
    auto none = [] { return GemmSeg<T>{}; };
    if (nn <= kLeaf) {
        GemmArgs<T> proto{};
        proto.skip_if = p.skip;
        proto.B = p.inv(j0);
        proto.ldb = kLeaf;
        proto.N = (int)nn;
        proto.K = (int)nn;
        proto.transb = 1;
        proto.alpha = T(1);
        proto.beta = T(0);
        if (int rc = over_groups<T>(G, proto, false, none,
                                    [&](const RowGroup<T>& g) {
                                        T* c = g.p + j0 * g.ld;
                                        return GemmSeg<T>{c, c, g.ld, g.ld, g.rows};
                                    },
                                    p.s))
            return rc;
        return p.sw ? split_cols_done(p, 0, j0, nn, false) : MOAO_OK;
    }
    const long n1 = split(nn);
    if (int rc = trsmT_rec(p, j0, n1)) return rc;
    // B[:, j0+n1:j0+nn] -= X[:, j0:j0+n1] * L[j0+n1:j0+nn, j0:j0+n1]^T
    GemmArgs<T> proto{};
    proto.skip_if = p.skip;
    proto.B = p.a(j0 + n1, j0);
    proto.ldb = p.lda;
    proto.N = (int)(nn - n1);
    proto.K = (int)n1;
    proto.transb = 1;
    proto.alpha = T(-1);
    proto.beta = T(1);
    if (p.sw) {
        proto.split = 2;
        proto.bp = p.sw->l_rows(j0 + n1, j0);
    }
    int rc = over_groups<T>(G, proto, false, none,
                            [&](const RowGroup<T>& g) {
                                GemmSeg<T> sg{g.p + j0 * g.ld, g.p + (j0 + n1) * g.ld, g.ld, g.ld,
                                              g.rows};
                                if (p.sw) sg.ap = p.sw->x_rows(j0);
                                return sg;
                            },
                            p.s);
    if (rc) return rc;
    return trsmT_rec(p, j0 + n1, nn - n1);
}

}  // namespace

__global__ void set_int_kernel(int* p, int v) { *p = v; }

template <typename T>
int potrf_tall(long n, T* A, long lda, const std::vector<RowGroup<T>>& groups, T* invL,
               int* info, hipStream_t s, const SplitWs* sw) {
    if (sw && groups.size() > 1) {
        set_error("split-BF16 planes take one row group");
        return MOAO_ERR_ARG;
    }
    Tall<T> p{n, A, lda, &groups, invL, info, info, s, sw};
    // a kernel rather than a memset node, so the sequence can be captured into a graph
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, info, 0);
    MOAO_HIP_TRY(hipGetLastError());
    return potrf_rec(p, 0, n);
}

template <typename T>
int trsm_right_lower_notrans(long n, const T* A, long lda, const T* invL,
                             const std::vector<RowGroup<T>>& groups, hipStream_t s,
                             const int* skip, const SplitWs* sw) {
    Tall<T> p{n, const_cast<T*>(A), lda, &groups, const_cast<T*>(invL), nullptr, skip, s, sw};
    return trsmN_rec(p, 0, n);
}

template <typename T>
int trsm_right_lower_trans(long n, const T* A, long lda, const T* invL,
                           const std::vector<RowGroup<T>>& groups, hipStream_t s,
                           const int* skip, const SplitWs* sw) {
    Tall<T> p{n, const_cast<T*>(A), lda, &groups, const_cast<T*>(invL), nullptr, skip, s, sw};
    return trsmT_rec(p, 0, n);
}

template int potrf_tall<double>(long, double*, long, const std::vector<RowGroup<double>>&,
                                double*, int*, hipStream_t, const SplitWs*);
template int potrf_tall<float>(long, float*, long, const std::vector<RowGroup<float>>&, float*,
                               int*, hipStream_t, const SplitWs*);
template int trsm_right_lower_notrans<double>(long, const double*, long, const double*,
                                              const std::vector<RowGroup<double>>&, hipStream_t,
                                              const int*, const SplitWs*);
template int trsm_right_lower_notrans<float>(long, const float*, long, const float*,
                                             const std::vector<RowGroup<float>>&, hipStream_t,
                                             const int*, const SplitWs*);
template int trsm_right_lower_trans<double>(long, const double*, long, const double*,
                                            const std::vector<RowGroup<double>>&, hipStream_t,
                                            const int*, const SplitWs*);
template int trsm_right_lower_trans<float>(long, const float*, long, const float*,
                                           const std::vector<RowGroup<float>>&, hipStream_t,
                                           const int*, const SplitWs*);

int reconstructor_f64_dev(long n, long m, double* dCmm, long ldCmm, double* dCtm, long ldCtm,
                          long ngal, double* dinvL, int* dinfo, hipStream_t stream) {
    if (n < 1 || m < 0 || ngal < 1 || ldCmm < n || (m > 0 && ldCtm < m) || !dCmm || !dinvL ||
        !dinfo || (m > 0 && !dCtm) || n >= (1L << 31) || m >= (1L << 31)) {
        set_error("reconstructor: invalid arguments");
        return MOAO_ERR_ARG;
    }
    // Target g's block starts at g*n*m (lapack/compute/reconstructor.cpp:43); with ldCtm > m
    // consecutive blocks would overlap and be updated concurrently.
    if (ngal > 1 && m > 0 && ldCtm != m) {
        set_error("reconstructor: several targets need ldCtm == nmeasts (blocks at g*n*m)");
        return MOAO_ERR_ARG;
    }
    std::vector<RowGroup<double>> groups;
    if (m > 0)
        for (long g = 0; g < ngal; ++g) groups.push_back({dCtm + g * n * m, ldCtm, (int)m});
    if (int rc = potrf_tall<double>(n, dCmm, ldCmm, groups, dinvL, dinfo, stream)) return rc;
    // the second solve is predicated on the status word: after a failing pivot it does nothing
    if (m > 0)
        if (int rc = trsm_right_lower_notrans<double>(n, dCmm, ldCmm, dinvL, groups, stream, dinfo))
            return rc;
    return MOAO_OK;
}

int reconstructor_f32_dev(long n, long m, float* dCmm, long ldCmm, float* dCtm, long ldCtm,
                          long ngal, float* dinvL, int* dinfo, hipStream_t stream) {
    if (n < 1 || m < 0 || ngal < 1 || ldCmm < n || (m > 0 && ldCtm < m) || !dCmm || !dinvL ||
        !dinfo || (m > 0 && !dCtm) || n >= (1L << 31) || m >= (1L << 31)) {
        set_error("reconstructor (FP32): invalid arguments");
        return MOAO_ERR_ARG;
    }
    if (ngal > 1 && m > 0 && ldCtm != m) {
        set_error("reconstructor (FP32): several targets need ldCtm == nmeasts");
        return MOAO_ERR_ARG;
    }
    std::vector<RowGroup<float>> groups;
    if (m > 0)
        for (long g = 0; g < ngal; ++g) groups.push_back({dCtm + g * n * m, ldCtm, (int)m});
    if (int rc = potrf_tall<float>(n, dCmm, ldCmm, groups, dinvL, dinfo, stream)) return rc;
    if (m > 0)
        if (int rc = trsm_right_lower_notrans<float>(n, dCmm, ldCmm, dinvL, groups, stream, dinfo))
            return rc;
    return MOAO_OK;
}

size_t recon_workspace_bytes(long n) {
    return ((recon_invl_doubles(n) * sizeof(double) + 255) & ~size_t(255)) + 256;
}

// ---- scratch of the allocation-free entry -------------------------------------------------
// The inverse leaves of one FP64 solve of order up to kNoAllocMaxN live in this array (40 MiB),
// which every device the library runs on holds until the process ends.  In host code the name
// `scratch` is only a host-side stand-in: kernels get the per-device address that
// hipGetSymbolAddress returns (taken once per device with that device current, at init or on
// the entry's first call there), never the symbol itself.
constexpr long kNoAllocMaxN = 40960;
constexpr size_t kNoAllocElems = (size_t)(kNoAllocMaxN / kLeaf) * kLeaf * kLeaf;
static __device__ __align__(256) double scratch[kNoAllocElems];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_setup() {
    int dev = -1;
    MOAO_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("allocation-free scratch: device index out of range");
        return MOAO_ERR_NODEV;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[dev]) return MOAO_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return MOAO_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double*>(p);
    return MOAO_OK;
}

static double* noalloc_scratch_ptr() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch[dev];
}

}  // namespace moao

extern "C" {

int moao_hip_reconstructor_async(long nmeas, long nmeasts, double* dCmm, long ldCmm,
                                 double* dCtm, long ldCtm, long ngal, int* d_info, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (!d_info || nmeas < 1) {
        moao::set_error("reconstructor_async: invalid arguments");
        return MOAO_ERR_ARG;
    }
    hipStream_t s = moao::pick_stream(stream);
    // workspace of this call only, from the stream-ordered pool: calls on different streams
    // never share it
    char* ws = nullptr;
    const size_t bytes = moao::recon_workspace_bytes(nmeas);
    if (hipMallocAsync(reinterpret_cast<void**>(&ws), bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        moao::set_error("reconstructor_async: workspace allocation failed");
        return MOAO_ERR_NOMEM;
    }
    double* invL = reinterpret_cast<double*>(ws);
    int* info = reinterpret_cast<int*>(ws + bytes - 256);
    int rc = moao::reconstructor_f64_dev(nmeas, nmeasts, dCmm, ldCmm, dCtm, ldCtm, ngal, invL,
                                         info, s);
    if (rc == MOAO_OK) {
        if (hipMemcpyAsync(d_info, info, sizeof(int), hipMemcpyDeviceToDevice, s) != hipSuccess) {
            moao::set_error("reconstructor_async: status copy failed");
            rc = MOAO_ERR_HIP;
        }
    }
    (void)hipFreeAsync(ws, s);
    return rc;
}

// The stream-ordered solve with its workspace in the static device scratch: no allocation,
// so callers that cannot allocate (graph capture, fixed serving loops) can use it.  *d_info is
// the solve's status word itself (cleared by a kernel at the start).
int moao_hip_reconstructor_noalloc(long nmeas, long nmeasts, double* dCmm, long ldCmm,
                                   double* dCtm, long ldCtm, long ngal, int* d_info,
                                   void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (!d_info || nmeas < 1) {
        moao::set_error("reconstructor_noalloc: invalid arguments");
        return MOAO_ERR_ARG;
    }
    if (nmeas > moao::kNoAllocMaxN) {
        moao::set_error("reconstructor_noalloc: nmeas = " + std::to_string(nmeas) +
                        " exceeds the scratch maximum of " + std::to_string(moao::kNoAllocMaxN) +
                        " (use moao_hip_reconstructor_async)");
        return MOAO_ERR_ARG;
    }
    // the device's scratch address (taken at init; here on the first call on another device)
    if (int rc = moao::noalloc_scratch_setup()) return rc;
    double* invL = moao::noalloc_scratch_ptr();
    if (!invL) {
        moao::set_error("reconstructor_noalloc: no scratch address for the current device");
        return MOAO_ERR_HIP;
    }
    return moao::reconstructor_f64_dev(nmeas, nmeasts, dCmm, ldCmm, dCtm, ldCtm, ngal, invL,
                                       d_info, moao::pick_stream(stream));
}

int moao_hip_reconstructor_dev(long nmeas, long nmeasts, double* dCmm, long ldCmm, double* dCtm,
                               long ldCtm, long ngal, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    hipStream_t s = moao::pick_stream(stream);
    int* dinfo = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&dinfo), sizeof(int), s) != hipSuccess) {
        (void)hipGetLastError();
        moao::set_error("reconstructor: status allocation failed");
        return MOAO_ERR_NOMEM;
    }
    int rc = moao_hip_reconstructor_async(nmeas, nmeasts, dCmm, ldCmm, dCtm, ldCtm, ngal, dinfo,
                                          s);
    int hinfo = 0;
    if (rc == MOAO_OK &&
        hipMemcpyAsync(&hinfo, dinfo, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) {
        moao::set_error("reconstructor: status copy failed");
        rc = MOAO_ERR_HIP;
    }
    (void)hipFreeAsync(dinfo, s);
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return rc != MOAO_OK ? rc : hinfo;
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
    // LAPACK's minimum extent of an n x n matrix with leading dimension ldCmm
    const size_t a_bytes = ((size_t)(nmeas - 1) * ldCmm + nmeas) * sizeof(double);
    // Ctm of target g starts at g*nmeas*nmeasts (lapack/compute/reconstructor.cpp:43).
    const size_t b_elems = nmeasts > 0 ? (size_t)(ngal - 1) * nmeas * nmeasts +
                                             (size_t)(nmeas - 1) * ldCtm + nmeasts
                                       : 0;
    const size_t ws_bytes = moao::recon_workspace_bytes(nmeas);
    if (int rc = c.host_stage.ensure(a_bytes)) return rc;
    if (int rc = c.host_stage2.ensure(b_elems * sizeof(double) + 8)) return rc;
    if (int rc = c.recon_ws.ensure(ws_bytes)) return rc;
    double* dA = c.host_stage.as<double>();
    double* dB = c.host_stage2.as<double>();
    double* invL = c.recon_ws.as<double>();
    int* info = reinterpret_cast<int*>(c.recon_ws.as<char>() + ws_bytes - 256);
    MOAO_HIP_TRY(hipMemcpyAsync(dA, Cmm, a_bytes, hipMemcpyHostToDevice, s));
    if (b_elems)
        MOAO_HIP_TRY(hipMemcpyAsync(dB, Ctm, b_elems * sizeof(double), hipMemcpyHostToDevice, s));
    int rc = moao::reconstructor_f64_dev(nmeas, nmeasts, dA, ldCmm, dB, ldCtm, ngal, invL, info, s);
    if (rc < 0) return rc;
    int hinfo = 0;
    MOAO_HIP_TRY(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    // Cmm gets the factor (on a failing pivot: every row of the columns of the leaves before
    // it, and the failing leaf's own rows of its 16-column blocks before the failing one; the
    // panel rows below a leaf are solved only once the leaf is whole); R only on success.
    MOAO_HIP_TRY(hipMemcpyAsync(Cmm, dA, a_bytes, hipMemcpyDeviceToHost, s));
    if (hinfo == 0 && b_elems)
        MOAO_HIP_TRY(hipMemcpyAsync(Ctm, dB, b_elems * sizeof(double), hipMemcpyDeviceToHost, s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return hinfo;
}

// FP32 storage: the Chameleon production precision (chameleon/moao_chameleon.cpp:41 runs
// MOAO_CHAMELEON::reconstructor<float>: spotrf + 2 strsm on the same maths).  Updates on FP32
// MFMA, leaves factored in FP64 and rounded to FP32.
int moao_hip_sreconstructor_dev(long nmeas, long nmeasts, float* dCmm, long ldCmm, float* dCtm,
                                long ldCtm, long ngal, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (nmeas < 1) {
        moao::set_error("sreconstructor: invalid arguments");
        return MOAO_ERR_ARG;
    }
    hipStream_t s = moao::pick_stream(stream);
    char* ws = nullptr;
    const size_t bytes = moao::recon_workspace_bytes(nmeas);  // doubles' room holds the floats
    if (hipMallocAsync(reinterpret_cast<void**>(&ws), bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        moao::set_error("sreconstructor: workspace allocation failed");
        return MOAO_ERR_NOMEM;
    }
    int* info = reinterpret_cast<int*>(ws + bytes - 256);
    int rc = moao::reconstructor_f32_dev(nmeas, nmeasts, dCmm, ldCmm, dCtm, ldCtm, ngal,
                                         reinterpret_cast<float*>(ws), info, s);
    int hinfo = 0;
    if (rc == MOAO_OK &&
        hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) {
        moao::set_error("sreconstructor: status copy failed");
        rc = MOAO_ERR_HIP;
    }
    (void)hipFreeAsync(ws, s);
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return rc != MOAO_OK ? rc : hinfo;
}

int moao_hip_sreconstructor(long nmeas, long nmeasts, float* Cmm, long ldCmm, float* Ctm,
                            long ldCtm, long ngal) {
    if (int rc = moao::ensure_init()) return rc;
    if (!Cmm || (nmeasts > 0 && !Ctm) || nmeas < 1 || nmeasts < 0 || ngal < 1 ||
        ldCmm < nmeas || (nmeasts > 0 && ldCtm < nmeasts)) {
        moao::set_error("sreconstructor: invalid arguments");
        return MOAO_ERR_ARG;
    }
    hipStream_t s = moao::ctx().stream;
    const size_t a_bytes = ((size_t)(nmeas - 1) * ldCmm + nmeas) * sizeof(float);
    const size_t b_elems = nmeasts > 0 ? (size_t)(ngal - 1) * nmeas * nmeasts +
                                             (size_t)(nmeas - 1) * ldCtm + nmeasts
                                       : 0;
    char* d = nullptr;
    const size_t a_al = (a_bytes + 255) & ~size_t(255);
    if (hipMallocAsync(reinterpret_cast<void**>(&d), a_al + b_elems * sizeof(float) + 256, s) !=
        hipSuccess) {
        (void)hipGetLastError();
        moao::set_error("sreconstructor: staging allocation failed");
        return MOAO_ERR_NOMEM;
    }
    float* dA = reinterpret_cast<float*>(d);
    float* dB = reinterpret_cast<float*>(d + a_al);
    int rc = MOAO_OK;
    if (hipMemcpyAsync(dA, Cmm, a_bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        (b_elems && hipMemcpyAsync(dB, Ctm, b_elems * sizeof(float), hipMemcpyHostToDevice, s) !=
                        hipSuccess)) {
        moao::set_error("sreconstructor: upload failed");
        rc = MOAO_ERR_HIP;
    }
    if (rc == MOAO_OK) rc = moao_hip_sreconstructor_dev(nmeas, nmeasts, dA, ldCmm, dB, ldCtm, ngal, s);
    if (rc >= 0) {
        // Cmm gets the (partial) factor; R only on success, as the FP64 entry
        if (hipMemcpyAsync(Cmm, dA, a_bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
            (rc == 0 && b_elems &&
             hipMemcpyAsync(Ctm, dB, b_elems * sizeof(float), hipMemcpyDeviceToHost, s) !=
                 hipSuccess)) {
            moao::set_error("sreconstructor: download failed");
            rc = MOAO_ERR_HIP;
        }
    }
    (void)hipFreeAsync(d, s);
    if (hipStreamSynchronize(s) != hipSuccess && rc >= 0) rc = MOAO_ERR_HIP;
    return rc;
}

}  // extern "C"
