"""GPU parity of the FP64 reconstructor with the reference's lapack path (the oracle's
scipy dpotrf + 2 x dtrsm in the reference's call order, oracle/lapack_path.py).

Tolerances.  Two LAPACKs on the same FP64 input already differ: on the reference's nx456
2-layer fixture (kappa(Cmm) ~ 2e7) OpenBLAS and MKL give R that differ by 1.2e-10 relative
Frobenius (backward errors ~1e-13 for both).  So R is checked two ways:
  * backward error  ||Ctm - R Cmm||_F / ||Ctm||_F <= 5e-13  (what a stable solver guarantees);
  * forward error   ||R - R_lapack||_F / ||R_lapack||_F <= 5e-10 on the kappa~2e7 fixtures and
    <= 1e-10 on the seeded small workload;
and the factor L (lower part) to 1e-12 relative Frobenius.
"""
import os

import numpy as np
import pytest

from conftest import FIXTURES, fixture_system
from moao_amd import matcov as gm
from moao_amd import systems
from moao_amd.reconstructor import reconstructor
from moao_amd.tomo import Tomo
from oracle import lapack_path as lp
from oracle.check import rel_fro

pytestmark = pytest.mark.gpu


def _inputs(sys_file, atm_file):
    t = Tomo(sys_file, atm_file)
    C = gm.cmm(t)
    t.set_target(0)
    B = gm.ctm(t)
    return C, B


def _sym(C):
    return np.tril(C) + np.tril(C, -1).T


def _check(C, B, fwd_tol):
    Cg, Bg = C.copy(order="F"), B.copy(order="F")
    n, m = C.shape[0], B.shape[0]
    info = reconstructor(n, m, Cg, n, Bg, m, 1)
    assert info == 0
    Co, Bo = C.copy(order="F"), B.copy(order="F")
    assert lp.reconstructor(Co, Bo) == 0
    bwd = np.linalg.norm(B - Bg @ _sym(C)) / np.linalg.norm(B)
    fwd = rel_fro(Bg, Bo)
    lfr = rel_fro(np.tril(Cg), np.tril(Co))
    print(f"    n={n} m={m} backward={bwd:.2e} forward={fwd:.2e} L={lfr:.2e}", flush=True)
    assert bwd <= 5e-13
    assert fwd <= fwd_tol
    assert lfr <= 1e-12
    # strictly-upper part of Cmm untouched (dpotrf 'L')
    iu = np.triu_indices(n, 1)
    assert np.array_equal(Cg[iu], C[iu])


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_systems(name):
    print(f"[recon] fixture {name}", flush=True)
    C, B = _inputs(*fixture_system(name))
    _check(C, B, 5e-10)


def test_small_workload(tmp_path):
    print("[recon] seeded small workload", flush=True)
    s, atms = systems.make_system(str(tmp_path), "small")
    C, B = _inputs(s, atms[0])
    _check(C, B, 1e-10)


def test_random_spd_sizes():
    rng = np.random.default_rng(3)
    for n, m in [(1, 3), (127, 5), (128, 128), (129, 40), (640, 257), (1500, 300)]:
        print(f"[recon] random SPD n={n} m={m}", flush=True)
        G = rng.standard_normal((n, n))
        C = np.asfortranarray(G @ G.T / n + np.eye(n))
        C[np.triu_indices(n, 1)] = 0.0
        B = np.asfortranarray(rng.standard_normal((m, n)))
        _check(C, B, 1e-11)


@pytest.mark.parametrize("n,m", [(9000, 1500), (20480, 2560)])
def test_large_sizes(n, m):
    """Sizes whose top-level updates launch >= 512 tiles of 128 rows: the trapezoid updates
    with the Ctm rows as a second segment and the K x N backward updates run on the 128-row
    kernels (the configuration of the n = 40k workload), checked like the small ones."""
    print(f"[recon] large random SPD n={n} m={m}", flush=True)
    rng = np.random.default_rng(n)
    # symmetric with a semicircle spectrum in [-1, 1], shifted: SPD with eigenvalues in [2, 4]
    G = rng.standard_normal((n, n))
    C = np.tril(G + G.T)
    del G
    C *= 0.25 / np.sqrt(n)
    C[np.diag_indices(n)] += 3.0
    C = np.asfortranarray(C)
    B = np.asfortranarray(rng.standard_normal((m, n)))
    _check(C, B, 1e-10)


def test_leading_dims_and_targets():
    """ldCmm > n, ldCtm > m and two targets stored n*m apart (reconstructor.cpp:43)."""
    print("[recon] leading dims + 2 targets", flush=True)
    rng = np.random.default_rng(5)
    n, m, ld_a, ld_b = 300, 40, 310, 40
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    Cbuf = np.zeros((ld_a, n), order="F")
    Cbuf[:n] = np.tril(S)
    B0, B1 = rng.standard_normal((m, n)), rng.standard_normal((m, n))
    Bbuf = np.concatenate([np.asfortranarray(B0).reshape(-1, order="F"),
                           np.asfortranarray(B1).reshape(-1, order="F")])
    info = reconstructor(n, m, Cbuf, ld_a, Bbuf, ld_b, 2)
    assert info == 0
    for g, Bg in enumerate((B0, B1)):
        R = Bbuf[g * n * m:(g + 1) * n * m].reshape((m, n), order="F")
        assert np.linalg.norm(Bg - R @ S) / np.linalg.norm(Bg) <= 5e-13
    assert np.all(Cbuf[n:] == 0.0)


def test_not_positive_definite_reports_info():
    print("[recon] non-SPD info", flush=True)
    rng = np.random.default_rng(9)
    n, m = 400, 10
    G = rng.standard_normal((n, n))
    C = np.asfortranarray(np.tril(G @ G.T / n + np.eye(n)))
    C[300, 300] = -1.0
    B = np.asfortranarray(rng.standard_normal((m, n)))
    Co = C.copy(order="F")
    info_ref = lp.reconstructor(Co, B.copy(order="F"))
    Bg = B.copy(order="F")
    info = reconstructor(n, m, C.copy(order="F"), n, Bg, m, 1)
    assert info == info_ref == 301
    assert np.array_equal(Bg, B)  # R not computed: Ctm untouched on the host entry


def test_nan_pivot_reports_info():
    """A NaN that reaches a pivot stops the factorization as a non-positive pivot does (LAPACK's
    dpotrf tests each pivot with `ajj <= 0 or disnan(ajj)`).  A NaN at (200, 150) leaves the
    pivots of columns 0..199 finite and reaches the pivot of column 200 first: info = 201, from
    inside the second 128-column leaf.  The oracle is not used here: OpenBLAS's own potrf does
    not test for NaN."""
    print("[recon] NaN pivot info", flush=True)
    rng = np.random.default_rng(11)
    n, m = 400, 10
    G = rng.standard_normal((n, n))
    C = np.asfortranarray(np.tril(G @ G.T / n + np.eye(n)))
    C[200, 150] = np.nan
    B = np.asfortranarray(rng.standard_normal((m, n)))
    Bg = B.copy(order="F")
    assert reconstructor(n, m, C.copy(order="F"), n, Bg, m, 1) == 201
    assert np.array_equal(Bg, B)
