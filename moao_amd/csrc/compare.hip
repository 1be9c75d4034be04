// compareMatrices (common/compute/check_utils.cpp:20-57): the reference's per-stage accuracy
// metric, on the host and on device-resident matrices.
//
//   maxErr = max over the visited entries of |1 - B/A|, or |B| where A == 0;
//   frobN  = sqrt(sum (B - A)^2), terms off the diagonal counted twice when uplo is 'U' or 'L'.
// A is column-major nrows x ncols (entry (i, j) at A[i + j*nrows]).  B's entry (i, j) is read
// at B[i + j*nrows] when trans == 'T' and at B[i*ncols + j] otherwise (the reference's own
// index rule: 'T' marks a B that a row-major FITS reader produced from a column-major
// writer).  uplo 'U' visits j >= i, 'L' visits j < i (the diagonal excluded, as the reference's
// loop bounds do), anything else every entry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace moao {
namespace {

constexpr int kCmpBlocks = 1024;

__device__ __forceinline__ void visit(const double* A, const double* B, long i, long j, long nrows,
                                      long ncols, bool tr, int ul, double& mx, double& fr) {
    const double a = A[i + j * nrows];
    const double b = tr ? B[i + j * nrows] : B[i * ncols + j];
    const double d = a == 0.0 ? -b : 1.0 - b / a;
    mx = fmax(mx, fabs(d));
    const double e = (b - a) * (b - a);
    fr += (ul == 0 || i == j) ? e : e * 2.0;
}

// Per-workgroup partial max and sum over the entries (column-major walk of A).
__global__ void compare_partial(const double* __restrict__ A, const double* __restrict__ B,
                                long nrows, long ncols, int tr, int ul, double* part) {
    __shared__ double smx[256], sfr[256];
    double mx = 0.0, fr = 0.0;
    const long total = nrows * ncols;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const long j = e / nrows, i = e - j * nrows;
        if (ul == 1 && j < i) continue;   // 'U': j >= i
        if (ul == 2 && j >= i) continue;  // 'L': j < i
        visit(A, B, i, j, nrows, ncols, tr != 0, ul, mx, fr);
    }
    smx[threadIdx.x] = mx;
    sfr[threadIdx.x] = fr;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + s]);
            sfr[threadIdx.x] += sfr[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smx[0];
        part[2 * blockIdx.x + 1] = sfr[0];
    }
}

int uplo_code(char uplo) { return uplo == 'U' ? 1 : (uplo == 'L' ? 2 : 0); }

}  // namespace
}  // namespace moao

extern "C" {

int moao_compare_matrices(const double* A, const double* B, long nrows, long ncols, char uplo,
                          char trans, double* max_err, double* frob) {
    if (!A || !B || nrows < 0 || ncols < 0 || !max_err || !frob) {
        moao::set_error("compare_matrices: invalid arguments");
        return MOAO_ERR_ARG;
    }
    const int ul = moao::uplo_code(uplo);
    double mx = 0.0, fr = 0.0;
    for (long i = 0; i < nrows; ++i) {
        const long start = ul == 1 ? i : 0, end = ul == 2 ? std::min(i, ncols) : ncols;
        for (long j = start; j < end; ++j) {
            const double a = A[i + j * nrows];
            const double b = trans == 'T' ? B[i + j * nrows] : B[i * ncols + j];
            const double d = a == 0.0 ? -b : 1.0 - b / a;
            mx = std::max(mx, std::fabs(d));
            const double e = (b - a) * (b - a);
            fr += (ul == 0 || i == j) ? e : e * 2.0;
        }
    }
    *max_err = mx;
    *frob = std::sqrt(fr);
    return MOAO_OK;
}

int moao_hip_compare_matrices_dev(const double* dA, const double* dB, long nrows, long ncols,
                                  char uplo, char trans, double* max_err, double* frob,
                                  void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    if (!dA || !dB || nrows < 0 || ncols < 0 || !max_err || !frob) {
        moao::set_error("compare_matrices_dev: invalid arguments");
        return MOAO_ERR_ARG;
    }
    hipStream_t s = moao::pick_stream(stream);
    double* part = nullptr;
    const size_t bytes = sizeof(double) * 2 * moao::kCmpBlocks;
    if (hipMallocAsync(reinterpret_cast<void**>(&part), bytes, s) != hipSuccess) {
        (void)hipGetLastError();
        moao::set_error("compare_matrices_dev: scratch allocation failed");
        return MOAO_ERR_NOMEM;
    }
    hipLaunchKernelGGL(moao::compare_partial, dim3(moao::kCmpBlocks), dim3(256), 0, s, dA, dB,
                       nrows, ncols, trans == 'T' ? 1 : 0, moao::uplo_code(uplo), part);
    std::vector<double> h(2 * moao::kCmpBlocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, bytes, hipMemcpyDeviceToHost, s);
    (void)hipFreeAsync(part, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        moao::set_error(std::string("compare_matrices_dev: ") + hipGetErrorString(e));
        return MOAO_ERR_HIP;
    }
    double mx = 0.0, fr = 0.0;
    for (int b = 0; b < moao::kCmpBlocks; ++b) {
        mx = std::max(mx, h[2 * b]);
        fr += h[2 * b + 1];
    }
    *max_err = mx;
    *frob = std::sqrt(fr);
    return MOAO_OK;
}

}  // extern "C"
