"""Pins the oracle's matcov restatement (CPU only).

1. Golden vectors from the hand-built reference (tests/golden/*.npz): sampled entries of Cmm,
   Ctm, Ctt must match bit for bit, and every column sum to 1e-15 relative.
2. Closed forms: the series coefficients are Gamma-function values, k1/k2/k3 and the Hankel
   coefficients have closed forms, and D(r) equals the exact von Karman structure function
   k1 L0^(5/3) (k2 - x^(5/6) K_{5/6}(x)) (x = 2 pi r / L0) built on scipy.special.kv, with the
   Kolmogorov limit 6.88 r^(5/3) as L0 -> infinity.
3. Invariants: Cmm symmetric positive definite; Ctm invariant under a common translation of
   every projected subaperture position; a layer above an LGS altitude contributes nothing.
"""
import math
import os

import numpy as np
import pytest
from scipy import special

from conftest import GOLDEN, SYSTEMS
from oracle import matcov as om
from oracle.tomo import Tomo

NAMES = sorted(n for n in os.listdir(SYSTEMS))


def _load(name):
    d = os.path.join(SYSTEMS, name)
    t = Tomo.from_files(os.path.join(d, "sys-params.txt"),
                        os.path.join(d, "prof-1-atmos-night0.txt"))
    t.set_target(0)
    return t, np.load(os.path.join(GOLDEN, name + ".npz"))


@pytest.mark.parametrize("name", NAMES)
def test_matrices_match_reference_golden(name):
    t, g = _load(name)
    mats = {"cmm": om.cmm(t), "ctm": om.ctm(t), "ctt": om.ctt(t)}
    for key, a in mats.items():
        i, j = g[key + "_idx"]
        assert np.array_equal(a[i, j], g[key + "_val"]), key
        full = np.tril(a) if key == "cmm" else a
        cs = full.sum(axis=0)
        assert np.max(np.abs(cs - g[key + "_colsum"])) <= 1e-15 * np.max(np.abs(cs)) + 1e-300
        assert abs(np.linalg.norm(full) - g[key + "_fro"][0]) <= 1e-15 * g[key + "_fro"][0]


# ---- closed forms of the constants (common/moao_defs.h:15-38) -----------------------------
A56 = 5.0 / 6.0
GA_REF = [0, 12.067619015983075, 5.17183672113560444, 0.795667187867016068,
          0.0628158306210802181, 0.00301515986981185091, 9.72632216068338833e-05,
          2.25320204494595251e-06, 3.93000356676612095e-08, 5.34694362825451923e-10,
          5.83302941264329804e-12]
GMA_REF = [-3.74878707653729304, -2.04479295083852408, -0.360845814853857083,
           -0.0313778969438136685, -0.001622994669507603, -5.56455315259749673e-05,
           -1.35720808599938951e-06, -2.47515152461894642e-08, -3.50257291219662472e-10,
           -3.95770950530691961e-12, -3.65327031259100284e-14]


def test_series_coefficients_are_gamma_values():
    for n in range(1, 11):
        ga = (-1) ** n * special.gamma(-n + A56) * 2 ** A56 / math.factorial(n)
        assert GA_REF[n] == pytest.approx(ga, rel=1e-14)
    for n in range(0, 11):
        gma = (-1) ** n * special.gamma(-n - A56) * 2 ** (-A56) / math.factorial(n)
        assert GMA_REF[n] == pytest.approx(gma, rel=1e-14)


def test_scalar_constants():
    k1 = (2 * special.gamma(11 / 6) * 2 ** (-5 / 6) * math.pi ** (-8 / 3)
          * (24 * special.gamma(6 / 5) / 5) ** (5 / 6))
    assert 0.1716613621245709486 == pytest.approx(k1, rel=1e-14)
    assert 1.00563491799858928388 == pytest.approx(special.gamma(5 / 6) * 2 ** (-1 / 6), rel=1e-14)
    assert 1.25331413731550012081 == pytest.approx(math.sqrt(math.pi / 2), rel=1e-15)
    # Hankel expansion of K_nu, mu = 4 nu^2 = 25/9: (mu-1)/8, (mu-1)(mu-9)/128, ...
    mu = 4 * A56 ** 2
    assert 0.22222222222222222222 == pytest.approx((mu - 1) / 8, rel=1e-15)
    assert -0.08641975308641974829 == pytest.approx((mu - 1) * (mu - 9) / 128, rel=1e-15)
    assert 0.08001828989483310284 == pytest.approx((mu - 1) * (mu - 9) * (mu - 25) / 3072,
                                                   rel=1e-15)
    lam2 = (0.5e-6 / (2 * math.pi) * 206265.0) ** 2
    assert 0.00026942094446267851 == pytest.approx(lam2, rel=1e-15)


def _dphi_exact(r, L0):
    k1 = 0.1716613621245709486
    k2 = special.gamma(5 / 6) * 2 ** (-1 / 6)
    x = 2 * math.pi * r / L0
    return k1 * L0 ** (5 / 3) * (k2 - x ** A56 * special.kv(A56, x))


@pytest.mark.parametrize("L0", [10.0, 25.0, 100.0])
def test_dphi_against_bessel_closed_form(L0):
    # the series (x <= 3 pi / 2) and the 3-term asymptote (x > 3 pi / 2) both truncate; the
    # largest relative error of the reference's model is 3.95e-6, at x = 3 pi / 2
    for r in np.concatenate([np.linspace(0.01, 30.0, 97), [L0 * 0.75 - 1e-9, L0 * 0.75 + 1e-9]]):
        ex = _dphi_exact(r, L0)
        got = om.dphi(r, 0.0, L0)
        assert got == pytest.approx(ex, rel=4e-6, abs=1e-12), (r, L0)
    assert om.dphi(0.0, 0.0, L0) == 0.0
    # isotropy: depends on |(x, y)| only
    assert om.dphi(3.0, 4.0, L0) == pytest.approx(om.dphi(5.0, 0.0, L0), rel=1e-15)


def test_dphi_kolmogorov_limit():
    # L0 -> infinity: D(r) -> 6.8839 r^(5/3) (r0 = 1)
    for r in (0.1, 0.5, 1.0, 3.0):
        assert om.dphi(r, 0.0, 1e12) == pytest.approx(6.88388 * r ** (5 / 3), rel=1e-3)


def test_cmm_spd_and_translation_invariance():
    t, _ = _load(NAMES[0])
    C = om.cmm(t)
    S = np.tril(C) + np.tril(C, -1).T
    assert np.linalg.eigvalsh(S).min() > 0
    B = om.ctm(t)
    t.u = t.u + 3.25
    t.v = t.v - 1.5
    B2 = om.ctm(t)
    assert np.max(np.abs(B2 - B)) <= 1e-12 * np.max(np.abs(B))


def test_layer_above_lgs_contributes_nothing():
    t, _ = _load(NAMES[0])
    L = t.atm.nLayer
    base = om.cmm(t)
    # move layer 0 above the LGS altitude (sspSizeL <= 0 for LGS) and zero its strength: the
    # LGS-LGS block must equal the one without that layer
    t.sspSizeL = t.sspSizeL.copy()
    for w in range(t.sys.nLgs):
        t.sspSizeL[w * L + 0] = -1.0
    c1 = om.cmm(t)
    t.cn2 = t.cn2.copy()
    t.cn2[0] = 0.0
    c2 = om.cmm(t)
    nl = 2 * sum(t.Nsubap[: t.sys.nLgs])
    assert np.array_equal(c1[:nl, :nl], c2[:nl, :nl])
    assert not np.array_equal(base[:nl, :nl], c1[:nl, :nl])
