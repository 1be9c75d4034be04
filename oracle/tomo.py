"""Tomo_struct preparation restated on the CPU (oracle; test infrastructure only).

Follows common/tomo_struct/tomo_struct.cpp (line numbers of that file):
  getSysParams        90-133   Nssp, pixSize, throughput, alpha/target in radians, sspSize, mr
  validSubap         379-471   annulus test (with the Nssp=7 / obs~0.286 special case), X, Y, Nsubap
  generateElongation 477-495   lgsExt, lgsTheta
  updateAtm          212-262   sspSizeL[w*L+l], L0diff (= L0 values), projection, noise
  setTargetCoord     291-303   truth-sensor direction, re-projection, re-noise
  updateSeeing       313-332   sNGS2, sLGS2 from r0 = (sum cn2)^(-3/5)
  updateNoise        338-374   noiseLGSxx/yy/xy[Nx], noiseNGS[nNgs+1]
  projectedSubap     501-565   u,v[l*Nx + pos]
  getNMeas/getNMeasTS 569-579
and common/tomo_struct/noise.cpp:16-31 (flux) / 36-52 (magnitudeFromFlux).

Scalar math uses Python's ``math`` module (the C library) in the reference's operation order,
so results match the reference to the last bit or ulp.

Pinned by: tests/test_oracle_tomo.py against golden vectors written by the hand-built
reference (tests/golden/make_golden.py) for the datafile/check nx456 fixtures and a seeded
synthetic system; plus invariants (Nx from the annulus rule, on-axis projection = X,Y at h=0).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .params import AtmParams, SysParams, read_atm, read_sys

RASC = 206265.0


def flux(qr, mr, throughput, d, t, bdw):
    """noise.cpp:16-31: e- per subaperture per frame."""
    return math.pow(10, qr - 0.4 * mr) * throughput * d * d * t * bdw * 1.0e4


def magnitude_from_flux(qr, F, throughput, d, t, bdw):
    """noise.cpp:36-52."""
    return -(math.log10(F / (throughput * d * d * bdw * 1.0e4 * t)) - qr) / 0.4


@dataclass
class Tomo:
    sys: SysParams
    atm: AtmParams

    # ---------------------------------------------------------------- construction
    @classmethod
    def from_files(cls, sys_file: str, atm_file: str) -> "Tomo":
        t = cls(read_sys(sys_file), AtmParams())
        t._get_sys_params()
        t._valid_subap()
        t._generate_elongation()
        t.update_atm(read_atm(atm_file))
        return t

    def _get_sys_params(self):
        s = self.sys
        self.qr = 2.9
        self.bdw = s.bdw_m * 1.0e10
        nW = s.nW
        self.Nssp = [int(s.nssp)] * nW
        self.pixSize = [s.lgsPixSize] * s.nLgs + [s.ngsPixSize] * (nW - s.nLgs)
        self.throughput = [s.throughLGS] * s.nLgs + [s.throughNGS] * (nW - s.nLgs)
        self.diamPup = [float(n) for n in self.Nssp]
        self.alphaX = [a * (1 / RASC) for a in s.alphaX_as[:nW]]
        self.alphaY = [a * (1 / RASC) for a in s.alphaY_as[:nW]]
        # targetX/Y are sized nW in the reference but filled from targetX_as (nTarget entries)
        self.targetX = [a * (1 / RASC) for a in s.targetX_as]
        self.targetY = [a * (1 / RASC) for a in s.targetY_as]
        self.sspSize = [s.diam / n for n in self.Nssp]
        mr = [0.0] * nW
        for i, v in enumerate(s.mrNGS[: nW - s.nLgs]):
            mr[s.nLgs + i] = v
        fpspf = (s.lgsFlux * s.tFrame * math.pow(s.diam / self.Nssp[0], 2.0) * self.throughput[0]
                 / s.throughAtm)
        for w in range(s.nLgs):
            mr[w] = magnitude_from_flux(self.qr, fpspf, self.throughput[w], self.sspSize[w],
                                        s.tFrame, self.bdw)
        self.mr = mr

    def _valid_subap(self):
        s = self.sys
        xs, ys, valid = [], [], []
        self.Nsubap = []
        for n in range(s.nW):
            nssp = self.Nssp[n]
            pas = s.diam / (1.0 * nssp)
            bornemin = -s.diam / 2.0
            rtel2 = (s.diam * s.diam) / 4.0
            if nssp != 7 or (s.obs <= 0.285 or s.obs >= 0.29):
                robs2 = s.diam * s.obs / 2.0 * s.diam * s.obs / 2.0
            else:
                robs2 = s.diam * 0.285 / 2.0 * s.diam * 0.285 / 2.0
            if nssp != 1:
                nsap = 0
                for i in range(nssp):
                    tp = bornemin + pas / 2.0 * (2.0 * i + 1.0)
                    for j in range(nssp):
                        x = bornemin + pas / 2.0 * (2.0 * j + 1.0)
                        r2 = x * x + tp * tp
                        ok = not (r2 < robs2 or r2 >= rtel2)
                        xs.append(x)
                        ys.append(tp)
                        valid.append(ok)
                        nsap += ok
                self.Nsubap.append(nsap)
            else:
                xs.append(0.0)
                ys.append(0.0)
                valid.append(True)
                self.Nsubap.append(1)
        self.X = np.array([x for x, ok in zip(xs, valid) if ok])
        self.Y = np.array([y for y, ok in zip(ys, valid) if ok])
        self.Nx = len(self.X)

    def _generate_elongation(self):
        s = self.sys
        self.lgsExt = np.zeros(self.Nx)
        self.lgsTheta = np.zeros(self.Nx)
        to_fwhm = RASC * s.lgsDepth / (s.lgsAlt * s.lgsAlt)
        k = 0
        for w in range(s.nLgs):
            ax, ay = self.alphaX[w], self.alphaY[w]
            den = math.sqrt(ax * ax + ay * ay)
            lltx = ax * (s.diam / 2.0 / den)
            llty = ay * (s.diam / 2.0 / den)
            for _ in range(self.Nsubap[w]):
                dx = self.X[k] - lltx
                dy = self.Y[k] - llty
                self.lgsExt[k] = math.sqrt(math.pow(dx, 2) + math.pow(dy, 2)) * to_fwhm
                self.lgsTheta[k] = math.atan2(dy * to_fwhm, dx * to_fwhm)
                k += 1

    # ---------------------------------------------------------------- atmosphere / target
    def update_atm(self, atm: AtmParams | str):
        if isinstance(atm, str):
            atm = read_atm(atm)
        self.atm = atm
        s, L = self.sys, atm.nLayer
        self.sspSizeL = np.zeros(s.nW * L)
        for cc in range(s.nW * L):
            n, l = divmod(cc, L)
            self.sspSizeL[cc] = self.sspSize[n] * (1.0 - s.gsAlt[n] * atm.h[l])
        self.L0diff = np.array(atm.L0[:L], dtype=float)
        self.cn2 = np.array(atm.cn2[:L], dtype=float)
        self._projected_subap()
        self._update_noise()

    def set_target(self, t: int):
        self.set_target_coord(self.targetX[t], self.targetY[t], 0)

    def set_target_coord(self, ax, ay, to_rad=0):
        if to_rad:
            ax /= RASC
            ay /= RASC
        self.alphaX[self.sys.nW - 1] = ax
        self.alphaY[self.sys.nW - 1] = ay
        self._projected_subap()
        self._update_noise()

    def _update_seeing(self):
        s = self.sys
        lam = 500.0 * 1.0e-9
        # the reference accumulates every entry of atm.cn2 (std::accumulate over the vector)
        acc = 0.0
        for c in self.atm.cn2:
            acc += c
        r0 = math.pow(acc, -3.0 / 5.0)
        v = lam / r0 * math.pow(lam / s.lambdaNGS, 0.2) * RASC
        self.sNGS2 = v * v
        v = lam / r0 * math.pow(lam / s.lambdaLGS, 0.2) * RASC
        self.sLGS2 = v * v

    def _update_noise(self):
        s = self.sys
        self.noiseNGS = np.zeros(s.nNgs + 1)
        self.noiseLGSxx = np.zeros(self.Nx)
        self.noiseLGSyy = np.zeros(self.Nx)
        self.noiseLGSxy = np.zeros(self.Nx)
        self._update_seeing()
        k = 0
        for i in range(s.nW):
            d = s.diam / self.Nssp[i]
            F = flux(self.qr, self.mr[i], self.throughput[i], d, s.tFrame, self.bdw)
            fact1 = 0.32 / F
            fact2 = 0.82 * math.pow(s.RON / (F * self.pixSize[i]), 2)
            if i < s.nLgs:
                pe = self.sLGS2 + s.spotWidth * s.spotWidth
                short = fact1 * pe + fact2 * math.pow(pe, 1.5) * math.pow(pe, 0.5)
                for _ in range(self.Nsubap[i]):
                    e2 = self.lgsExt[k] * self.lgsExt[k]
                    long_ = (fact1 * (pe + e2) + fact2 * math.pow(pe + e2, 1.5) * math.pow(pe, 0.5))
                    th = self.lgsTheta[k]
                    c, sn = math.cos(th), math.sin(th)
                    self.noiseLGSxx[k] = long_ * math.pow(c, 2) + short * math.pow(sn, 2)
                    self.noiseLGSyy[k] = long_ * math.pow(sn, 2) + short * math.pow(c, 2)
                    self.noiseLGSxy[k] = (long_ - short) * sn * c
                    k += 1
            else:
                self.noiseNGS[i - s.nLgs] = fact1 * self.sNGS2 + fact2 * self.sNGS2 * self.sNGS2

    def _projected_subap(self):
        s, L, Nx = self.sys, self.atm.nLayer, self.Nx
        rad = 3.14159265358979323846 / 180.0
        ioff = [0] * s.nW
        for i in range(1, s.nW):
            ioff[i] = ioff[i - 1] + self.Nsubap[i - 1]
        self.u = np.zeros(L * Nx)
        self.v = np.zeros(L * Nx)
        wfs_of = np.zeros(Nx, dtype=int)
        for n in range(s.nW):
            wfs_of[ioff[n]: ioff[n] + self.Nsubap[n]] = n
        for l in range(L):
            h = self.atm.h[l]
            for pos in range(Nx):
                n = int(wfs_of[pos])
                i = pos - ioff[n]
                dX = self.alphaX[n] * h
                dY = self.alphaY[n] * h
                rr = 1.0 - h * s.gsAlt[n]
                G = self.diamPup[n] / float(self.Nssp[n])
                th = s.thetaML[n] * rad
                xtp = self.X[ioff[n] + i] * G
                ytp = self.Y[ioff[n] + i] * G
                uu = xtp * math.cos(th) - ytp * math.sin(th)
                vv = xtp * math.sin(th) + ytp * math.cos(th)
                uu += s.XPup[n]
                vv += s.YPup[n]
                self.u[l * Nx + pos] = uu * rr + dX
                self.v[l * Nx + pos] = vv * rr + dY

    # ---------------------------------------------------------------- sizes
    def nmeas(self) -> int:
        return 2 * (self.Nx - self.Nsubap[self.sys.nW - 1])

    def nmeasts(self) -> int:
        return 2 * self.Nsubap[self.sys.nLgs + self.sys.nNgs]
