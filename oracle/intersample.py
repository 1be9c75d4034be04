"""intersample_process restated with numpy FFTs in FP64 (oracle; test infrastructure only).

Follows common/tomo_struct/intersample.cpp:72-171 step by step:
  Map[idx[i]-1] += Cvv[i] (in index order), Map /= div where div > 0          83-95
  map[j + i*N] = Map[cpt] on the dactupix lattice centred at ncmap = N/2+1     98-111
  r2c FFT, tmp.re *= abs2fi[j*N+i], tmp.im = 0, c2r                           113-123
  norm = dphi[N*(ncmap-1)+ncmap-1] fact/fact2; dphi = exp(-0.5(norm - dphi fact/fact2))
  with fact = 2 (2*3.14159/lambda)^2, fact2 = N^2 dactupix^2                   131-143
  _eclat; tmp.re = otf_tel[j*N+i]*dphi[j*N+i], tmp.im = 0; c2r; _eclat          145-167
numpy.fft.rfft2 / irfft2 are FFTW's r2c / c2r transforms (irfft2 * N^2 = FFTW's unnormalised
c2r); _eclat (intersample.cpp:47-70) on an even N is fftshift on both axes.
N = max(512, 8*2^floor(ln(np)+0.5)) and dactupix = N/(3(np-1)) + 1 (integer division), 175-207.

Pinned by tests/test_oracle_intersample.py against what the mathematics fixes, on inputs that
are not symmetric under a transpose (so a swapped axis anywhere fails):
  * a stationary anisotropic Cvv_ab = g(x_a - x_b, y_a - y_b): every map bin equals g at its
    offset (index orientation, scatter, hit-count division);
  * white Cvv and a separable Gaussian Cvv with an anisotropic abs2fi: dphi equals the closed
    form exp(-fact/(2 fact2) (raw(0) - raw)) with raw from direct-sum autocorrelations
    (fact, fact2, dactupix, centring, abs2fi's [j*N+i] orientation);
  * PSF pixels by a direct O(N^2) cosine sum with an elliptical OTF, Strehl = sum(otf * dphi);
  * a literal transcription of the lattice-embedding loop and of the _eclat loops;
  * the half-spectrum pipeline against full complex FFTs, and Cvv = 0 -> Strehl = sum(otf) = 1.
"""
from __future__ import annotations

import math

import numpy as np

LAMBDA = 1.65e-06


def default_n(np_: int) -> int:
    return int(max(512, 8 * 2 ** int(math.log(np_) + 0.5)))


def dactupix(N: int, np_: int) -> float:
    return float(N // (3 * (np_ - 1)) + 1)


def eclat(a: np.ndarray) -> np.ndarray:
    return np.fft.fftshift(a)


def cvv_map(cvv_flat, idx, np_):
    ncov = 2 * np_ - 1
    nmap = ncov * ncov
    Map = np.zeros(nmap)
    np.add.at(Map, idx - 1, cvv_flat)
    div = np.bincount(idx - 1, minlength=nmap).astype(float)
    nz = div > 0
    Map[nz] /= div[nz]
    return Map


def embed(Map, N, np_):
    ncov = 2 * np_ - 1
    dp = dactupix(N, np_)
    ncmap = N // 2 + 1
    nelem = (ncov - 1) // 2
    first = int(ncmap - nelem * dp - 1)
    pos = first + np.arange(ncov) * int(dp)
    m = np.zeros((N, N))
    m[np.ix_(pos, pos)] = Map.reshape(ncov, ncov)  # m[i][j] = map[j + i*N]
    return m


def process(cvv_flat, idx, abs2fi_flat, otf_flat, N, np_, return_dphi=False):
    dp = dactupix(N, np_)
    m = embed(cvv_map(np.asarray(cvv_flat, float), np.asarray(idx), np_), N, np_)
    h = N // 2 + 1
    T = np.fft.rfft2(m)
    a2f = np.asarray(abs2fi_flat, float).reshape(N, N)   # a2f[j][i] = abs2fi[j*N + i]
    T = T.real * a2f.T[:, :h]
    dphi = np.fft.irfft2(T, s=(N, N)) * (N * N)
    fact = 2 * (2 * 3.14159 / LAMBDA) ** 2
    fact2 = N * N * dp * dp
    norm = dphi[N // 2, N // 2] / fact2 * fact
    dphi = np.exp(-0.5 * (norm - dphi / fact2 * fact))
    dphi = eclat(dphi)
    otf = np.asarray(otf_flat, float).reshape(N, N)
    T2 = (otf * dphi).T[:, :h]
    psf = eclat(np.fft.irfft2(T2, s=(N, N)) * (N * N))
    return (psf, dphi) if return_dphi else psf
