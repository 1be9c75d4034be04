// Batched 2-D complex FFTs for the PSF stage (power-of-two N, 32 <= N <= 4096).
#pragma once

#include <hip/hip_runtime.h>

namespace moao {

template <typename R>
struct Cplx;
template <>
struct Cplx<float> { using type = float2; };
template <>
struct Cplx<double> { using type = double2; };

// Twiddle table W_N^e = exp(-2 pi i e / N), e < N, built on the host in FP64.
template <typename R>
int fft_twiddles(int N, typename Cplx<R>::type** d_table, hipStream_t s);

// In-place 2-D FFT of `batch` row-major N x N complex arrays in X (scratch: same size).
// dir = -1 forward (FFTW_FORWARD sign), +1 backward; unnormalised like FFTW.
template <typename R>
int fft2d(typename Cplx<R>::type* X, typename Cplx<R>::type* scratch, int N, int batch, int dir,
          const typename Cplx<R>::type* tw, hipStream_t s);

// The same transform without its final transpose: the result of X goes to Y transposed
// (Y[j][i] = F[i][j]); X is overwritten.  Since the 2-D transform commutes with transposition,
// a transposed input gives the plain-layout result.
template <typename R>
int fft2d_t(typename Cplx<R>::type* X, typename Cplx<R>::type* Y, int N, int batch, int dir,
            const typename Cplx<R>::type* tw, hipStream_t s);

}  // namespace moao
