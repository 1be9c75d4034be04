// Intersample PSF (common/tomo_struct/intersample.cpp:72-171) on gfx950, batched over
// directions, FP32 (or FP64) FFTs.
//
// Reference steps and their counterparts here (per direction):
//   Map[idx[i]-1] += Cvv[i]; Map /= div          map_gather: one wave per map bin over a CSR
//                                                 list of the Cvv entries falling in it
//                                                 (deterministic, no atomics)
//   embed Map in the N x N grid at stride dactupix embed (after a zero fill), two directions
//                                                 per complex grid (real and imaginary parts)
//   r2c FFT; *= abs2fi[j*N+i] (real part); c2r    complex FFTs of the packed pair (each real
//                                                 transform costs half a complex one); the two
//                                                 spectra separated by Hermitian symmetry, the
//                                                 half-spectrum product extended to the full
//                                                 grid by the mirror the c2r transform assumes;
//                                                 transforms skip their final transpose, the
//                                                 spectra are held transposed
//   dphi = exp(-0.5 (norm - dphi/fact2*fact))     dphi_kernel (fftshift folded in)
//   _eclat; tmp = otf_tel[j*N+i]*dphi[j*N+i]; c2r spec2 (mirror as above), backward FFT
//   _eclat -> PSF                                 final_kernel (fftshift folded in)
// with the reference's constants: lambda = 1.65e-6, fact = 2 (2*3.14159/lambda)^2,
// fact2 = N^2 dactupix^2, dactupix = N/(3(np-1)) + 1 (integer division), norm at
// (N/2, N/2).  N defaults to the reference's max(512, 8*2^floor(ln(np)+0.5))
// (intersample.cpp:192) and may be set explicitly (a 2048 grid is not reachable by that rule).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "profile.hpp"

struct moao_isample {
    long N = 0, np = 0, nidx = 0, ncov = 0, nmap = 0;
    double lambda = 1.65e-06, dactupix = 0;
    int prec = 0;  // 0 FP32, 1 FP64
    int batch_cap = 0;
    int* bin_off = nullptr;    // [nmap+1] CSR offsets
    int* bin_ent = nullptr;    // [nidx] Cvv flat indices grouped by bin
    double* div = nullptr;     // [nmap] hit counts
    void* a2f = nullptr;       // [N*N] abs2fi (R), reference flat layout
    void* otf = nullptr;       // [N*N] otf_tel (R)
    void* tw = nullptr;        // twiddles
    void* X = nullptr;         // [batch][N*N] complex
    void* Y = nullptr;
    double* Map = nullptr;     // [batch][nmap]
    double* dCvv = nullptr;    // host-entry staging
    double* dPsf = nullptr;
};

namespace moao {
namespace {

// Workgroup order.  Entry k of bin b + 1 is usually the Cvv element next to entry k of bin b
// (the same actuator pair shifted by one actuator), so consecutive bins read the same cache
// lines.  The dispatcher deals workgroups to the 8 XCDs round robin; launch index L is
// re-mapped so that each XCD's runs of kGatherRun workgroups take 4 * kGatherRun consecutive
// bins, which then meet in that XCD's L2, while the XCDs still share every part of the map
// (a bijection on [0, workgroups): indices past the last whole round keep their place).
constexpr int kGatherRun = 16;

__global__ void map_gather(const double* __restrict__ cvv, long nidx, const int* __restrict__ off,
                           const int* __restrict__ ent, const double* __restrict__ div,
                           double* __restrict__ Map, long nmap) {
    long L = blockIdx.x + (long)gridDim.x * blockIdx.y;
    const long total = (long)gridDim.x * gridDim.y, round = 8L * kGatherRun;
    if (L < total / round * round) {
        const long c = L % 8, q = L / 8;
        L = (q / kGatherRun) * round + c * kGatherRun + q % kGatherRun;
    }
    const long bin = (L % gridDim.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int b = (int)(L / gridDim.x);
    if (bin >= nmap) return;
    const double* c = cvv + (size_t)b * nidx;
    double acc = 0.0;
    for (int e = off[bin] + lane; e < off[bin + 1]; e += 64) acc += c[ent[e]];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) {
        const double d = div[bin];
        Map[(size_t)b * nmap + bin] = d > 0 ? acc / d : acc;
    }
}

template <typename C>
__global__ void zero_fill(C* X, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) X[i] = C{0, 0};
}

// Two directions per complex grid: direction 2p in the real part, 2p + 1 in the imaginary part
// (zero when the batch is odd).  Every transform of the PSF is a transform of real data, so one
// complex transform carries two of them: the forward spectra separate by Hermitian symmetry and
// the later spectra are real-valued and symmetric, so their inverse transforms are real.
template <typename C>
__global__ void embed(const double* __restrict__ Map, long nmap, C* X, int N, int ncov, int first,
                      int step, int batch) {
    const int t = blockIdx.x * 256 + threadIdx.x;  // map index = ti * ncov + tj
    const int p = blockIdx.y;                      // direction pair
    if (t >= ncov * ncov) return;
    const int ti = t / ncov, tj = t % ncov;
    const int i = first + ti * step, j = first + tj * step;
    using R = decltype(C{}.x);
    const R re = (R)Map[(size_t)(2 * p) * nmap + t];
    const R im = 2 * p + 1 < batch ? (R)Map[(size_t)(2 * p + 1) * nmap + t] : R(0);
    X[(size_t)p * N * N + (size_t)i * N + j] = C{re, im};
}

// G(i, j) = (Ga, Gb) with G(i, j) = Re F(i', j') * w[j'*N + i'], (i', j') = (i, j) for
// j <= N/2, else the Hermitian mirror ((N-i)%N, N-j): the full-grid form of the half spectrum a
// c2r transform reads.  F_a = (Z(k) + conj Z(-k)) / 2 and F_b = (Z(k) - conj Z(-k)) / 2i of the
// packed transform Z, so Re F_a = (Re Z(k) + Re Z(-k)) / 2, Re F_b = (Im Z(k) + Im Z(-k)) / 2.
// Both Z and G are held transposed (Z(a, b) at X[b*N + a], G(i, j) at Y[j*N + i]), the layout
// the transforms leave and take without their final transpose; lanes run along i, so every
// read (Z, its mirror, w[j'*N + i']) and the write are contiguous.
template <typename R, typename C>
__global__ void spec_mul(const C* __restrict__ X, const R* __restrict__ w, C* __restrict__ Y,
                         int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    const int p = blockIdx.z;
    if (i >= N) return;
    int ii = i, jj = j;
    if (j > N / 2) {
        ii = (N - i) % N;
        jj = N - j;
    }
    const size_t o = (size_t)p * N * N;
    const C z = X[o + (size_t)jj * N + ii];
    const C zm = X[o + (size_t)((N - jj) % N) * N + (N - ii) % N];
    R wk = w[(size_t)jj * N + ii];
    // columns 0 and N/2 hold both k and -k: the weight is averaged over the pair, so the
    // product stays even whatever the weights (a c2r transform keeps only that even part)
    if (jj == 0 || jj == N / 2) wk = (R)0.5 * (wk + w[(size_t)jj * N + (N - ii) % N]);
    Y[o + (size_t)j * N + i] = C{(R)0.5 * (z.x + zm.x) * wk, (R)0.5 * (z.y + zm.y) * wk};
}

// dphi = exp(-0.5 (norm - D fact / fact2)), norm = D(N/2, N/2) fact / fact2, then fftshift;
// D of direction 2p in the real part of Y, of 2p + 1 in the imaginary part; stored packed in X.
template <typename R, typename C>
__global__ void dphi_kernel(const C* __restrict__ Y, C* __restrict__ X, int N, double fact,
                            double fact2) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    const int p = blockIdx.z;
    if (j >= N) return;
    const size_t o = (size_t)p * N * N;
    const C c = Y[o + (size_t)(N / 2) * N + N / 2];
    const int si = (i + N / 2) % N, sj = (j + N / 2) % N;  // value landing at (i, j) after shift
    const C d = Y[o + (size_t)si * N + sj];
    const double na = (double)c.x / fact2 * fact, nb = (double)c.y / fact2 * fact;
    X[o + (size_t)i * N + j] = C{(R)exp(-0.5 * (na - (double)d.x / fact2 * fact)),
                                 (R)exp(-0.5 * (nb - (double)d.y / fact2 * fact))};
}

// spec2: G(i, j) = otf[j'*N+i'] * dphi[j'*N+i'] with the same mirror for j > N/2, both
// directions of the pair; dphi as dphi_kernel stores it, G written transposed (at Y[j*N + i])
// with lanes along i, as in spec_mul.
template <typename R, typename C>
__global__ void spec_otf(const C* __restrict__ X, const R* __restrict__ otf, C* __restrict__ Y,
                         int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    const int p = blockIdx.z;
    if (i >= N) return;
    int ii = i, jj = j;
    if (j > N / 2) {
        ii = (N - i) % N;
        jj = N - j;
    }
    const size_t o = (size_t)p * N * N;
    const size_t k = (size_t)jj * N + ii;  // flat index j'*N + i'
    const C d = X[o + k];
    if (jj == 0 || jj == N / 2) {  // even part, as in spec_mul
        const size_t km = (size_t)jj * N + (N - ii) % N;
        const C dm = X[o + km];
        Y[o + (size_t)j * N + i] = C{(R)0.5 * (otf[k] * d.x + otf[km] * dm.x),
                                     (R)0.5 * (otf[k] * d.y + otf[km] * dm.y)};
        return;
    }
    Y[o + (size_t)j * N + i] = C{otf[k] * d.x, otf[k] * d.y};
}

template <typename C>
__global__ void final_kernel(const C* __restrict__ Y, double* __restrict__ psf, int N, int batch) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    const int p = blockIdx.z;
    if (j >= N) return;
    const size_t o = (size_t)p * N * N;
    const int si = (i + N / 2) % N, sj = (j + N / 2) % N;
    const C v = Y[o + (size_t)si * N + sj];
    psf[(size_t)(2 * p) * N * N + (size_t)i * N + j] = (double)v.x;
    if (2 * p + 1 < batch) psf[(size_t)(2 * p + 1) * N * N + (size_t)i * N + j] = (double)v.y;
}

// Long-exposure sum: sum[i] += psf[i] (one FP64 addition per element and step, so the sum
// equals a host loop that adds the steps' PSFs in the same order).
__global__ void psf_accumulate(const double* __restrict__ psf, long count,
                               double* __restrict__ sum) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) sum[i] += psf[i];
}

template <typename R>
int process(moao_isample* S, const double* dCvv, int batch, double* dPsf, hipStream_t s) {
    using C = typename Cplx<R>::type;
    const int N = (int)S->N, ncov = (int)S->ncov;
    const int pairs = (batch + 1) / 2;  // complex grids: two directions each
    if (batch > S->batch_cap) {
        const size_t cb = sizeof(C) * (size_t)N * N * pairs;
        if (S->X) (void)hipFree(S->X);
        if (S->Y) (void)hipFree(S->Y);
        if (S->Map) (void)hipFree(S->Map);
        S->X = S->Y = nullptr;
        S->Map = nullptr;
        S->batch_cap = 0;
        MOAO_HIP_TRY(hipMalloc(&S->X, cb));
        MOAO_HIP_TRY(hipMalloc(&S->Y, cb));
        MOAO_HIP_TRY(hipMalloc((void**)&S->Map, sizeof(double) * S->nmap * batch));
        S->batch_cap = batch;
    }
    C* X = (C*)S->X;
    C* Y = (C*)S->Y;
    const C* tw = (const C*)S->tw;
    const size_t tot = (size_t)N * N * pairs;
    {
        ProfScope ps(kProfOther, 8.0 * S->nidx * batch, s);
        hipLaunchKernelGGL(map_gather, dim3((S->nmap + 3) / 4, batch), dim3(256), 0, s, dCvv,
                           S->nidx, S->bin_off, S->bin_ent, S->div, S->Map, S->nmap);
        hipLaunchKernelGGL(zero_fill<C>, dim3((tot + 255) / 256), dim3(256), 0, s, X, tot);
        const int nelem = (ncov - 1) / 2, ncmap = N / 2 + 1, step = (int)S->dactupix;
        const int first = (int)(ncmap - nelem * S->dactupix - 1);
        hipLaunchKernelGGL(embed<C>, dim3((ncov * ncov + 255) / 256, pairs), dim3(256), 0, s,
                           S->Map, S->nmap, X, N, ncov, first, step, batch);
        MOAO_HIP_TRY(hipGetLastError());
    }
    // Transforms without their final transpose: a plain-layout input comes out transposed and
    // a transposed one plain, so the spectra of steps 1 and 3 are held transposed and D and
    // the PSF plain.
    if (int rc = fft2d_t<R>(X, Y, N, pairs, -1, tw, s)) return rc;  // Y: Z transposed
    const dim3 pg((N + 255) / 256, N, pairs);
    hipLaunchKernelGGL((spec_mul<R, C>), pg, dim3(256), 0, s, Y, (const R*)S->a2f, X, N);
    if (int rc = fft2d_t<R>(X, Y, N, pairs, +1, tw, s)) return rc;  // Y: D
    const double fact = 2 * std::pow(2 * 3.14159 / S->lambda, 2);
    const double fact2 = (double)N * N * S->dactupix * S->dactupix;
    hipLaunchKernelGGL((dphi_kernel<R, C>), pg, dim3(256), 0, s, Y, X, N, fact, fact2);
    hipLaunchKernelGGL((spec_otf<R, C>), pg, dim3(256), 0, s, X, (const R*)S->otf, Y, N);
    if (int rc = fft2d_t<R>(Y, X, N, pairs, +1, tw, s)) return rc;  // X: the PSF before shift
    hipLaunchKernelGGL(final_kernel<C>, pg, dim3(256), 0, s, X, dPsf, N, batch);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

}  // namespace
}  // namespace moao

using moao::ctx;
using moao::set_error;

extern "C" {

long moao_intersample_default_n(long np) {
    // intersample.cpp:192: MAX(512, 8*pow(2, (long)(log(np)+0.5)))  (natural log)
    const double v = 8 * std::pow(2.0, (double)(long)(std::log((double)np) + 0.5));
    return (long)std::max(512.0, v);
}

int moao_hip_intersample_create(long nidx, long np, long N, const long* idx, const double* abs2fi,
                                const double* otf_tel, int precision, moao_isample** out) {
    if (int rc = moao::ensure_init()) return rc;
    if (!out || !idx || !abs2fi || !otf_tel || nidx < 1 || np < 2 || precision < 0 ||
        precision > 1) {
        set_error("intersample_create: invalid arguments");
        return MOAO_ERR_ARG;
    }
    if (N <= 0) N = moao_intersample_default_n(np);
    if (N < 32 || N > 4096 || (N & (N - 1))) {
        set_error("intersample_create: N must be a power of two in [32, 4096]");
        return MOAO_ERR_ARG;
    }
    auto* S = new moao_isample;
    S->N = N;
    S->np = np;
    S->nidx = nidx;
    S->ncov = 2 * np - 1;
    S->nmap = S->ncov * S->ncov;
    S->prec = precision;
    S->dactupix = (double)(N / (3 * (np - 1)) + 1);  // intersample.cpp:199 (integer division)
    const long nelem = (S->ncov - 1) / 2, ncmap = N / 2 + 1;
    const long first = (long)(ncmap - nelem * S->dactupix - 1);
    const long last = (long)(first + (S->ncov - 1) * S->dactupix);
    if (first < 0 || last >= N) {
        delete S;
        set_error("intersample_create: the covariance map does not fit the N x N grid");
        return MOAO_ERR_ARG;
    }
    // CSR of Cvv entries per map bin (intersample_init computes div the same way).
    std::vector<int> cnt(S->nmap + 1, 0);
    for (long i = 0; i < nidx; ++i) {
        const long b = idx[i] - 1;
        if (b < 0 || b >= S->nmap) {
            delete S;
            set_error("intersample_create: idx entry out of range");
            return MOAO_ERR_ARG;
        }
        cnt[b + 1]++;
    }
    std::vector<double> div(S->nmap);
    for (long b = 0; b < S->nmap; ++b) div[b] = cnt[b + 1];
    for (long b = 0; b < S->nmap; ++b) cnt[b + 1] += cnt[b];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1), ent(nidx);
    for (long i = 0; i < nidx; ++i) ent[fill[idx[i] - 1]++] = (int)i;
    hipStream_t s = ctx().stream;
    auto fail = [&](int rc) {
        moao_hip_intersample_destroy(S);
        return rc;
    };
    auto H2D = [&](void** d, const void* h, size_t bytes) -> int {
        if (hipMalloc(d, bytes) != hipSuccess) return MOAO_ERR_NOMEM;
        if (hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return MOAO_ERR_HIP;
        return MOAO_OK;
    };
    if (int rc = H2D((void**)&S->bin_off, cnt.data(), sizeof(int) * cnt.size())) return fail(rc);
    if (int rc = H2D((void**)&S->bin_ent, ent.data(), sizeof(int) * nidx)) return fail(rc);
    if (int rc = H2D((void**)&S->div, div.data(), sizeof(double) * S->nmap)) return fail(rc);
    const size_t nn = (size_t)N * N;
    if (precision == 0) {
        std::vector<float> a(nn), o(nn);
        for (size_t k = 0; k < nn; ++k) a[k] = (float)abs2fi[k], o[k] = (float)otf_tel[k];
        if (int rc = H2D(&S->a2f, a.data(), sizeof(float) * nn)) return fail(rc);
        if (int rc = H2D(&S->otf, o.data(), sizeof(float) * nn)) return fail(rc);
        if (hipStreamSynchronize(s) != hipSuccess) return fail(MOAO_ERR_HIP);
        if (int rc = moao::fft_twiddles<float>((int)N, (float2**)&S->tw, s)) return fail(rc);
    } else {
        if (int rc = H2D(&S->a2f, abs2fi, sizeof(double) * nn)) return fail(rc);
        if (int rc = H2D(&S->otf, otf_tel, sizeof(double) * nn)) return fail(rc);
        if (int rc = moao::fft_twiddles<double>((int)N, (double2**)&S->tw, s)) return fail(rc);
    }
    if (hipStreamSynchronize(s) != hipSuccess) return fail(MOAO_ERR_HIP);
    *out = S;
    return MOAO_OK;
}

long moao_hip_intersample_n(const moao_isample* S) { return S ? S->N : MOAO_ERR_ARG; }

int moao_hip_psf_accumulate_dev(const double* dPsf, long count, double* dSum, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (count < 0 || (count > 0 && (!dPsf || !dSum))) {
        set_error("psf_accumulate: invalid arguments");
        return MOAO_ERR_ARG;
    }
    if (count == 0) return MOAO_OK;
    hipStream_t s = moao::pick_stream(stream);
    hipLaunchKernelGGL(moao::psf_accumulate, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       s, dPsf, count, dSum);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

int moao_hip_intersample_process_dev(moao_isample* S, const double* dCvv, int batch,
                                     double* dPsf, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (!S || !dCvv || !dPsf || batch < 1) {
        set_error("intersample_process: invalid arguments");
        return MOAO_ERR_ARG;
    }
    hipStream_t s = moao::pick_stream(stream);
    return S->prec == 0 ? moao::process<float>(S, dCvv, batch, dPsf, s)
                        : moao::process<double>(S, dCvv, batch, dPsf, s);
}

// Host drop-in of intersample_process (intersample.cpp:72-171): Cvv is the flat nact x nact
// covariance (nidx doubles), psf receives the N x N PSF in the reference's layout.
int moao_hip_intersample_process(moao_isample* S, const double* Cvv, double* psf) {
    if (int rc = moao::ensure_init()) return rc;
    if (!S || !Cvv || !psf) return MOAO_ERR_ARG;
    hipStream_t s = ctx().stream;
    if (!S->dCvv) MOAO_HIP_TRY(hipMalloc((void**)&S->dCvv, sizeof(double) * S->nidx));
    if (!S->dPsf) MOAO_HIP_TRY(hipMalloc((void**)&S->dPsf, sizeof(double) * S->N * S->N));
    MOAO_HIP_TRY(hipMemcpyAsync(S->dCvv, Cvv, sizeof(double) * S->nidx, hipMemcpyHostToDevice, s));
    if (int rc = moao_hip_intersample_process_dev(S, S->dCvv, 1, S->dPsf, s)) return rc;
    MOAO_HIP_TRY(hipMemcpyAsync(psf, S->dPsf, sizeof(double) * S->N * S->N, hipMemcpyDeviceToHost,
                                s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return MOAO_OK;
}

void moao_hip_intersample_destroy(moao_isample* S) {
    if (!S) return;
    void* ptrs[] = {S->bin_off, S->bin_ent, S->div, S->a2f, S->otf, S->tw, S->X, S->Y, S->Map,
                    S->dCvv, S->dPsf};
    (void)hipDeviceSynchronize();
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete S;
}

}  // extern "C"
