// Batched 2-D complex FFT for gfx950: LDS-resident Stockham row passes + tiled transposes.
//
// A row of N complex values lives in LDS (ping-pong buffers); log2(N) is covered by radix-4
// Stockham stages (one radix-2 stage first when log2(N) is odd), one workgroup per row, N/4
// lanes each doing one 4-point butterfly per stage, so a row costs one HBM read and one write.
// The 2-D transform is rows -> transpose -> rows -> transpose (32x32 LDS tiles, padded).
// Twiddles come from a table computed on the host in FP64 and rounded once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "profile.hpp"

namespace moao {
namespace {

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
// multiply by -i*dir... : for forward (dir=-1) the radix-4 kernel needs x * (-i); backward x * (+i)
template <typename C>
__device__ __forceinline__ C mul_i(C a, int dir) {  // a * (dir * i)
    return dir > 0 ? C{-a.y, a.x} : C{a.y, -a.x};
}

template <typename R, int NMAX>
__global__ void fft_rows(typename Cplx<R>::type* X, int N, int log2N, int dir,
                         const typename Cplx<R>::type* __restrict__ tw) {
    using C = typename Cplx<R>::type;
    __shared__ C buf[2][NMAX];
    C* row = X + (size_t)blockIdx.x * N;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < N; e += nt) buf[0][e] = row[e];
    __syncthreads();
    int cur = 0, Ns = 1;
    if (log2N & 1) {  // radix-2 stage
        const int h = N / 2;
        for (int j = tid; j < h; j += nt) {
            const C a = buf[cur][j], b = buf[cur][j + h];
            const int base = 2 * j;  // Ns == 1: (j - k) * 2 + k with k = 0
            buf[cur ^ 1][base] = cadd(a, b);
            buf[cur ^ 1][base + 1] = csub(a, b);
        }
        __syncthreads();
        cur ^= 1;
        Ns = 2;
    }
    const int q = N / 4;
    while (Ns < N) {
        const int stride = N / (Ns * 4);  // twiddle exponent step
        for (int j = tid; j < q; j += nt) {
            const int k = j & (Ns - 1);
            C v0 = buf[cur][j], v1 = buf[cur][j + q], v2 = buf[cur][j + 2 * q],
              v3 = buf[cur][j + 3 * q];
            if (Ns > 1) {
                C w1 = tw[k * stride], w2 = tw[2 * k * stride], w3 = tw[3 * k * stride];
                if (dir > 0) {
                    w1.y = -w1.y;
                    w2.y = -w2.y;
                    w3.y = -w3.y;
                }
                v1 = cmul(v1, w1);
                v2 = cmul(v2, w2);
                v3 = cmul(v3, w3);
            }
            const C s02 = cadd(v0, v2), d02 = csub(v0, v2);
            const C s13 = cadd(v1, v3), d13 = mul_i(csub(v1, v3), dir);
            const int base = (j - k) * 4 + k;
            buf[cur ^ 1][base] = cadd(s02, s13);
            buf[cur ^ 1][base + Ns] = cadd(d02, d13);
            buf[cur ^ 1][base + 2 * Ns] = csub(s02, s13);
            buf[cur ^ 1][base + 3 * Ns] = csub(d02, d13);
        }
        __syncthreads();
        cur ^= 1;
        Ns *= 4;
    }
    for (int e = tid; e < N; e += nt) row[e] = buf[cur][e];
}

template <typename C>
__global__ void transpose(const C* __restrict__ in, C* __restrict__ out, int N) {
    __shared__ C tile[32][33];
    const size_t off = (size_t)blockIdx.z * N * N;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[ty + 8 * r][tx] = in[off + (size_t)(y0 + ty + 8 * r) * N + x0 + tx];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) out[off + (size_t)(x0 + ty + 8 * r) * N + y0 + tx] = tile[tx][ty + 8 * r];
}

template <typename R>
int rows(typename Cplx<R>::type* X, int N, int nrows, int dir, const typename Cplx<R>::type* tw,
         hipStream_t s) {
    int log2N = 0;
    while ((1 << log2N) < N) ++log2N;
    const int nt = std::max(64, N / 4);
    if (N <= 512)
        hipLaunchKernelGGL((fft_rows<R, 512>), dim3(nrows), dim3(nt), 0, s, X, N, log2N, dir, tw);
    else if (N <= 1024)
        hipLaunchKernelGGL((fft_rows<R, 1024>), dim3(nrows), dim3(nt), 0, s, X, N, log2N, dir, tw);
    else if (N <= 2048)
        hipLaunchKernelGGL((fft_rows<R, 2048>), dim3(nrows), dim3(nt), 0, s, X, N, log2N, dir, tw);
    else
        hipLaunchKernelGGL((fft_rows<R, 4096>), dim3(nrows), dim3(nt), 0, s, X, N, log2N, dir, tw);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

}  // namespace

template <typename R>
int fft_twiddles(int N, typename Cplx<R>::type** d_table, hipStream_t s) {
    using C = typename Cplx<R>::type;
    std::vector<C> h(N);
    const double pi = 3.14159265358979323846;
    for (int e = 0; e < N; ++e) {
        const double a = -2.0 * pi * (double)e / (double)N;
        h[e] = C{(R)std::cos(a), (R)std::sin(a)};
    }
    MOAO_HIP_TRY(hipMalloc((void**)d_table, sizeof(C) * N));
    MOAO_HIP_TRY(hipMemcpyAsync(*d_table, h.data(), sizeof(C) * N, hipMemcpyHostToDevice, s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return MOAO_OK;
}

template <typename R>
int fft2d(typename Cplx<R>::type* X, typename Cplx<R>::type* Y, int N, int batch, int dir,
          const typename Cplx<R>::type* tw, hipStream_t s) {
    using C = typename Cplx<R>::type;
    if (N < 32 || N > 4096 || (N & (N - 1)) || batch < 1) {
        set_error("fft2d: N must be a power of two in [32, 4096]");
        return MOAO_ERR_ARG;
    }
    ProfScope ps(kProfFft, 4.0 * 2.0 * sizeof(C) * (double)N * N * batch, s);
    const dim3 tg(N / 32, N / 32, batch);
    if (int rc = rows<R>(X, N, N * batch, dir, tw, s)) return rc;
    hipLaunchKernelGGL(transpose<C>, tg, dim3(256), 0, s, X, Y, N);
    if (int rc = rows<R>(Y, N, N * batch, dir, tw, s)) return rc;
    hipLaunchKernelGGL(transpose<C>, tg, dim3(256), 0, s, Y, X, N);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

template <typename R>
int fft2d_t(typename Cplx<R>::type* X, typename Cplx<R>::type* Y, int N, int batch, int dir,
            const typename Cplx<R>::type* tw, hipStream_t s) {
    using C = typename Cplx<R>::type;
    if (N < 32 || N > 4096 || (N & (N - 1)) || batch < 1) {
        set_error("fft2d: N must be a power of two in [32, 4096]");
        return MOAO_ERR_ARG;
    }
    ProfScope ps(kProfFft, 3.0 * 2.0 * sizeof(C) * (double)N * N * batch, s);
    const dim3 tg(N / 32, N / 32, batch);
    if (int rc = rows<R>(X, N, N * batch, dir, tw, s)) return rc;
    hipLaunchKernelGGL(transpose<C>, tg, dim3(256), 0, s, X, Y, N);
    MOAO_HIP_TRY(hipGetLastError());
    return rows<R>(Y, N, N * batch, dir, tw, s);
}

template int fft_twiddles<float>(int, float2**, hipStream_t);
template int fft_twiddles<double>(int, double2**, hipStream_t);
template int fft2d<float>(float2*, float2*, int, int, int, const float2*, hipStream_t);
template int fft2d<double>(double2*, double2*, int, int, int, const double2*, hipStream_t);
template int fft2d_t<float>(float2*, float2*, int, int, int, const float2*, hipStream_t);
template int fft2d_t<double>(double2*, double2*, int, int, int, const double2*, hipStream_t);

}  // namespace moao
