"""Pins the oracle's LAPACK-path restatement (CPU only): textbook routines, closed forms and
the in-place geadd mirror rule of the reference."""
import numpy as np
import pytest
import scipy.linalg as sl

from conftest import FIXTURES, fixture_system
from oracle import lapack_path as lp
from oracle import matcov as om
from oracle.tomo import Tomo


def _system(name):
    t = Tomo.from_files(*fixture_system(name))
    C = om.cmm(t)
    t.set_target(0)
    return C, om.ctm(t), om.ctt(t)


def _sym(C):
    return np.tril(C) + np.tril(C, -1).T


@pytest.mark.parametrize("name", FIXTURES)
def test_reconstructor_solves_the_system(name):
    C, B, _ = _system(name)
    S = _sym(C)
    Cw, R = C.copy(order="F"), B.copy(order="F")
    assert lp.reconstructor(Cw, R) == 0
    assert np.linalg.norm(B - R @ S) / np.linalg.norm(B) < 5e-13
    R2 = sl.cho_solve(sl.cho_factor(S, lower=True), B.T).T
    assert np.linalg.norm(R - R2) / np.linalg.norm(R2) < 1e-9
    # L is the Cholesky factor of S
    Lf = np.tril(Cw)
    assert np.linalg.norm(Lf @ Lf.T - S) / np.linalg.norm(S) < 1e-14


def test_reconstructor_reports_info():
    rng = np.random.default_rng(1)
    G = rng.standard_normal((50, 50))
    C = np.asfortranarray(np.tril(G @ G.T + 50 * np.eye(50)))
    C[20, 20] = -1.0
    B = np.asfortranarray(rng.standard_normal((4, 50)))
    B0 = B.copy()
    assert lp.reconstructor(C, B) == 21
    assert np.array_equal(B, B0)


def test_geadd_mirror_rule():
    # laset(U,0) + lacpy(L) leaves [[a,0,0],[b,c,0],[d,e,f]]; the in-place geadd('T') loop
    # makes it symmetric with a doubled diagonal
    B = np.asfortranarray(np.array([[1.0, 0, 0], [2.0, 3.0, 0], [4.0, 5.0, 6.0]]))
    lp.geadd_T_inplace(B)
    assert np.array_equal(B, np.array([[2.0, 2, 4], [2, 6, 5], [4, 5, 12]]))


@pytest.mark.parametrize("name", FIXTURES[:1])
def test_cee_cvv_closed_form(name):
    C, Ctm, Ctt = _system(name)
    S = _sym(C)
    R = Ctm.copy(order="F")
    lp.reconstructor(C.copy(order="F"), R)
    m = Ctt.shape[0]
    nact = 60
    rng = np.random.default_rng(2)
    Dx = np.asfortranarray(rng.standard_normal((nact, m)) / np.sqrt(C.shape[0]))
    Cpp, Cpm = Ctt.copy(order="F"), Ctm.copy(order="F")
    Cee, Cvv = lp.cee_cvv(C, Cpp, Cpm, R, Dx)
    ref = Ctt - Ctm @ R.T - R @ Ctm.T + R @ S @ R.T
    assert np.linalg.norm(Cee - ref) / np.linalg.norm(ref) < 1e-9
    # Cee is a small difference of large terms (|Cee| ~ 1e-3 |Ctt|) and the final dgemm is not
    # symmetric in its rounding, so symmetry holds to ~1e-11 relative
    assert np.linalg.norm(Cee - Cee.T) / np.linalg.norm(Cee) < 1e-10
    assert np.linalg.norm(Cvv - Dx @ ref @ Dx.T) / np.linalg.norm(Cvv) < 1e-9
    # side effects of the reference: Cpm holds R Cmm, Cpp's lower part the syr2k result
    assert np.linalg.norm(Cpm - R @ S) / np.linalg.norm(Cpm) < 1e-13
    low = np.tril_indices(m)
    want = (Ctt - Ctm @ R.T - R @ Ctm.T)[low]
    assert np.linalg.norm(Cpp[low] - want) / np.linalg.norm(want) < 1e-12
    # the literal geadd loop gives the same Cee as the vectorised mirror
    Cee2, _ = lp.cee_cvv(C, Ctt.copy(order="F"), Ctm.copy(order="F"), R, Dx, small_geadd=True)
    assert np.array_equal(Cee, Cee2)
