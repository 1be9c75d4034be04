"""Seeded inputs of the intersample PSF stage: DM geometry, Cvv-to-map index, |FT(influence)|^2
and the telescope OTF, as the reference's input pipeline defines them.

  make_dm        cython/pipeline/commande.py:57-79   (Ndiam = nssp+1 actuators across, valid
                 where r < 1 + 1.4/nssp and r > obs - 1/nssp, x0 = sqrt(-2/ln coupling)/nssp)
  make_idx       cython/pipeline/generate_file.py:57-76  (1-based index of the actuator-pair
                 offset (dx, dy) in the (2 Ndiam - 1)^2 map; Cvv flat index a + b*nact)
  make_abs2fi    generate_file.py:79-86 with funcInflu (commande.py:160-174): |FFT2(fi)|^2 of
                 the Gaussian influence function sampled at ud = 3 D / N
  make_otf_tel   cython/pipeline/intersample.pyx:174-200: autocorrelation of the annular pupil
                 normalised so the diffraction-limited PSF peak (Strehl) is 1
  interaction_matrix / command_matrix
                 cython/pipeline/commande.py:81-127: the truth sensor's interaction matrix from
                 the influence function and Dx as its truncated-SVD pseudo-inverse
Arrays are returned flat in the reference's (FITS) memory order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass
class DM:
    ndiam: int
    csI: np.ndarray   # flat indices (row-major in an ndiam x ndiam image) of valid actuators
    x0: float         # influence-function width, pupil-radius units

    @property
    def nact(self) -> int:
        return int(self.csI.size)


def make_dm(nssp: int, coupling: float = 0.2, obs: float = 0.28) -> DM:
    nd = nssp + 1
    x = np.tile(np.linspace(-1, 1, nd), (nd, 1))
    y = x.T.copy()
    r = np.sqrt(x * x + y * y).flatten()
    nn = np.where((r < (1.0 + 1.4 / nssp)) & (r > (obs - 1.0 / nssp)))[0]
    return DM(nd, nn, math.sqrt(-2.0 / math.log(coupling)) / nssp)


def make_idx(dm: DM) -> np.ndarray:
    Np = dm.ndiam
    xs = dm.csI % Np           # column of each actuator
    ys = dm.csI // Np          # row
    dx = xs[:, None] - xs[None, :] + Np
    dy = ys[:, None] - ys[None, :] + Np
    ind = dx + (2 * Np - 1) * (dy - 1)      # [a][b]
    return np.asarray(ind.flatten(order="F"), dtype=np.int64)  # index a + b*nact


def grid_step(N: int, diam: float) -> float:
    """ud: the phase-map field is 3 telescope diameters across N pixels."""
    return 3.0 * diam / N


def make_abs2fi(N: int, diam: float, dm: DM) -> np.ndarray:
    ud = grid_step(N, diam)
    c = (np.arange(N) + 1 - (N // 2 + 1)) * ud / (diam / 2.0)
    y = np.tile(c, (N, 1)).T
    x = y.T
    fi = 1.0e-6 * np.exp(-(x * x + y * y) / (2 * dm.x0 * dm.x0))
    a = np.abs(np.fft.fft2(fi.T)) ** 2
    return np.ascontiguousarray(a).flatten()


def make_otf_tel(N: int, diam: float, obs: float) -> np.ndarray:
    ud = grid_step(N, diam)
    x = ud / (diam / 2.0) * (np.arange(N) + 1 - (N // 2 + 1))
    x2 = np.tile(x * x, (N, 1))
    r = np.sqrt(x2 + x2.T)
    pup = ((r <= 1.0) & (r > obs)).astype(np.float64)
    surface_pix = diam ** 2 * (1 - obs ** 2) * np.pi / 4.0 / ud ** 2
    b = np.abs(np.fft.fft2(pup))
    ac = np.fft.ifft2(b * b).real
    otf = ac / surface_pix ** 2
    return np.ascontiguousarray(otf.T).flatten()


def actuator_positions(dm: DM):
    """(csX, csY) of the valid actuators in pupil-radius units (commande.py:61-70)."""
    lin = np.linspace(-1, 1, dm.ndiam)
    return lin[dm.csI % dm.ndiam], lin[dm.csI // dm.ndiam]


def interaction_matrix(X: np.ndarray, Y: np.ndarray, nssp: int, diam: float, dm: DM) -> np.ndarray:
    """Interaction matrix of the truth sensor (commande.py:81-106): slopes (x slopes of every
    valid subaperture, then y slopes; arcsec) per volt of each actuator, from the difference of
    the Gaussian influence function (funcInflu, commande.py:160-174) across each subaperture.
    X, Y: the sensor's valid subaperture centres in metres, in the system's order."""
    sx, sy = np.asarray(X) / (diam / 2.0), np.asarray(Y) / (diam / 2.0)
    cx, cy = actuator_positions(dm)
    dx = sx[None, :] - cx[:, None]          # actuator x subaperture
    dy = sy[None, :] - cy[:, None]
    ss = (2.0 / nssp) / 2.0

    def fi(x, y):
        return 1.0e-6 * np.exp(-(x * x + y * y) / (2 * dm.x0 * dm.x0))

    nsub = sx.size
    mia = np.zeros((2 * nsub, dm.nact))
    mia[:nsub, :] = (fi(dx + ss, dy) - fi(dx - ss, dy)).T
    mia[nsub:, :] = (fi(dy + ss, dx) - fi(dy - ss, dx)).T
    return mia / (diam / nssp) * 206265.0


def command_matrix(mia: np.ndarray, thresh: float = 0.0, condi: float = 0.0) -> np.ndarray:
    """Dx = the truncated-SVD pseudo-inverse of the interaction matrix (computeCommandMat,
    commande.py:108-127): singular values below thresh (max/sqrt(condi) when condi is set,
    1e-4 when neither is) are filtered.  Shape nact x 2 nsub, column-major, FP64."""
    U, w, Vt = np.linalg.svd(mia, full_matrices=False)
    if condi != 0.0:
        thresh = np.max(w) / np.sqrt(condi)
    if thresh == 0.0:
        thresh = 1e-4
    winv = np.where(w < thresh, 0.0, 1.0 / np.where(w < thresh, 1.0, w))
    return np.asfortranarray((Vt.T * winv) @ U.T)
