"""GPU parity of Cee/Cvv with the reference's lapack chain (oracle/lapack_path.cee_cvv:
scipy dsyr2k, dsymm, laset/lacpy/geadd/scal, dgemm, dsymm, dgemm in the reference's order).

Tolerances: Cee is a small difference of large terms (|Cee| ~ 1e-3 |Ctt| on these systems), so
its relative error is the FP64 rounding of the O(|Ctt|) terms divided by |Cee|: 1e-10 relative
Frobenius; Cvv = Dx Cee Dx^T inherits it (1e-10); the side effects (Cpm = R Cmm, lower Cpp) are
O(1) quantities: 1e-12.
"""
import numpy as np
import pytest

from conftest import FIXTURES, fixture_system
from moao_amd import matcov as gm
from moao_amd.cee_cvv import cee_cvv
from moao_amd.systems import dx_matrix
from moao_amd.tomo import Tomo
from oracle import lapack_path as lp
from oracle.check import rel_fro

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", FIXTURES)
def test_cee_cvv_matches_lapack_chain(name):
    print(f"[cee] fixture {name}", flush=True)
    t = Tomo(*fixture_system(name))
    C = gm.cmm(t)
    t.set_target(0)
    Ctm, Ctt = gm.ctm(t), gm.ctt(t)
    n, m = C.shape[0], Ctt.shape[0]
    R = Ctm.copy(order="F")
    lp.reconstructor(C.copy(order="F"), R)
    nact = 137
    Dx = dx_matrix(nact, m, n)
    # oracle
    Cpp_o, Cpm_o = Ctt.copy(order="F"), Ctm.copy(order="F")
    Cee_o, Cvv_o = lp.cee_cvv(C, Cpp_o, Cpm_o, R, Dx)
    # GPU
    Cpp, Cpm = Ctt.copy(order="F"), Ctm.copy(order="F")
    Cee = np.zeros((m, m), order="F")
    Cvv = np.zeros((nact, nact), order="F")
    Tmp = np.zeros((nact, m), order="F")
    cee_cvv(n, m, nact, C, Cpp, Cpm, R, Dx, Cee, Cvv, Tmp)
    e_cee, e_cvv = rel_fro(Cee, Cee_o), rel_fro(Cvv, Cvv_o)
    e_cpm = rel_fro(Cpm, Cpm_o)
    low = np.tril_indices(m)
    e_cpp = rel_fro(Cpp[low], Cpp_o[low])
    print(f"    Cee {e_cee:.2e} Cvv {e_cvv:.2e} Cpm {e_cpm:.2e} Cpp {e_cpp:.2e}", flush=True)
    assert e_cee <= 1e-10 and e_cvv <= 1e-10
    assert e_cpm <= 1e-12 and e_cpp <= 1e-12
    # upper part of Cpp keeps the caller's values (dsyr2k 'L')
    up = np.triu_indices(m, 1)
    assert np.array_equal(Cpp[up], Ctt[up])
    assert np.linalg.norm(Tmp - Dx @ Cee_o) / np.linalg.norm(Tmp) <= 1e-10


def test_symmetric_b_gemm_mode():
    """transb=2 reads only the lower triangle of B (upper filled with garbage)."""
    from moao_amd.device import DeviceArray
    from moao_amd.reconstructor import dgemm_dev
    rng = np.random.default_rng(4)
    for M, K in [(50, 300), (257, 129), (7, 16)]:
        print(f"[cee] symmetric-B gemm M={M} K={K}", flush=True)
        S = rng.standard_normal((K, K))
        S = S + S.T
        B = np.tril(S) + np.triu(np.full((K, K), 1e300), 1)
        A = rng.standard_normal((M, K))
        dA, dB = DeviceArray.from_host(A), DeviceArray.from_host(np.asfortranarray(B))
        dC = DeviceArray.from_host(np.zeros((M, K), order="F"))
        dgemm_dev(2, M, K, K, 1.0, dA.ptr, M, dB.ptr, K, 0.0, dC.ptr, M)
        ref = A @ S
        assert np.max(np.abs(dC.to_host() - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_cee_cvv_large_synthetic():
    """n = 9000: step 2 (R sym(Cmm), 12 x 71 tiles of 128 rows) runs on the symmetric
    direct-to-LDS kernel with Cmm's full diagonal-block copies; the whole chain is checked
    against the oracle on synthetic operands (random SPD Cmm with garbage above its diagonal,
    random symmetric Ctt, random Ctm and R: no cancellation, so FP64 rounding alone, 1e-11)."""
    n, m, nact = 9000, 1500, 600
    print(f"[cee] large synthetic n={n} m={m} nact={nact}", flush=True)
    rng = np.random.default_rng(21)
    G = rng.standard_normal((n, n))
    C = np.tril(G + G.T) * (0.25 / np.sqrt(n))
    del G
    C[np.diag_indices(n)] += 3.0
    C[np.triu_indices(n, 1)] = 1e300  # never read (dsymm 'L')
    C = np.asfortranarray(C)
    T = rng.standard_normal((m, m))
    Ctt = np.asfortranarray(T + T.T)
    Ctm = np.asfortranarray(rng.standard_normal((m, n)))
    R = np.asfortranarray(rng.standard_normal((m, n)) * 0.01)
    Dx = np.asfortranarray(rng.standard_normal((nact, m)))
    Cl = np.asfortranarray(np.tril(C))
    Cpp_o, Cpm_o = Ctt.copy(order="F"), Ctm.copy(order="F")
    Cee_o, Cvv_o = lp.cee_cvv(Cl, Cpp_o, Cpm_o, R, Dx)
    Cpp, Cpm = Ctt.copy(order="F"), Ctm.copy(order="F")
    Cee = np.zeros((m, m), order="F")
    Cvv = np.zeros((nact, nact), order="F")
    Tmp = np.zeros((nact, m), order="F")
    cee_cvv(n, m, nact, C, Cpp, Cpm, R, Dx, Cee, Cvv, Tmp)
    e_cee, e_cvv, e_cpm = rel_fro(Cee, Cee_o), rel_fro(Cvv, Cvv_o), rel_fro(Cpm, Cpm_o)
    print(f"    Cee {e_cee:.2e} Cvv {e_cvv:.2e} Cpm {e_cpm:.2e}", flush=True)
    assert e_cee <= 1e-11 and e_cvv <= 1e-11 and e_cpm <= 1e-12
    assert np.all(np.isfinite(Cee)) and np.all(np.isfinite(Cvv))


@pytest.mark.parametrize("ngal", [3, 9])
def test_batched_targets_equal_per_target_calls(ngal):
    """moao_hip_cee_cvv_batched_dev (every target's products in shared launches, 8 targets per
    launch) against one moao_hip_cee_cvv_dev call per target: every output and side effect
    identical bit for bit, ragged sizes, 9 targets = two launches per step."""
    import ctypes
    from moao_amd.device import DeviceArray
    from moao_amd.lib import check, lib
    print(f"[cee] batched ngal={ngal}", flush=True)
    n, m, nact = 700, 301, 137
    rng = np.random.default_rng(ngal)
    G = rng.standard_normal((n, n))
    Cmm = np.asfortranarray(G @ G.T / n + np.eye(n))
    Cmm[np.triu_indices(n, 1)] = 1e300  # only the lower triangle may be read
    Dx = np.asfortranarray(rng.standard_normal((nact, m)))
    Ctt = [rng.standard_normal((m, m)) for _ in range(ngal)]
    Ctt = [np.asfortranarray(a + a.T) for a in Ctt]
    Ctm = [np.asfortranarray(rng.standard_normal((m, n))) for _ in range(ngal)]
    R = [np.asfortranarray(rng.standard_normal((m, n))) for _ in range(ngal)]
    vp = ctypes.c_void_p
    dCmm, dDx = DeviceArray.from_host(Cmm), DeviceArray.from_host(Dx)
    # batched: target g's blocks side by side along the columns
    b = {"Cpp": DeviceArray.from_host(np.hstack(Ctt)), "Cpm": DeviceArray.from_host(np.hstack(Ctm)),
         "R": DeviceArray.from_host(np.hstack(R)), "Cee": DeviceArray((m, m * ngal)),
         "Cvv": DeviceArray((nact, nact * ngal)), "Tmp": DeviceArray((nact, m * ngal))}
    check(lib().moao_hip_cee_cvv_batched_dev(
        n, m, nact, ngal, vp(dCmm.ptr), n, vp(b["Cpp"].ptr), m, vp(b["Cpm"].ptr), m,
        vp(b["R"].ptr), m, vp(dDx.ptr), nact, vp(b["Cee"].ptr), m, vp(b["Cvv"].ptr), nact,
        vp(b["Tmp"].ptr), nact, None), "cee_cvv_batched")
    got = {k: v.to_host() for k, v in b.items()}
    width = {"Cpp": m, "Cpm": n, "R": n, "Cee": m, "Cvv": nact, "Tmp": m}
    for g in range(ngal):
        s = {"Cpp": DeviceArray.from_host(Ctt[g]), "Cpm": DeviceArray.from_host(Ctm[g]),
             "R": DeviceArray.from_host(R[g]), "Cee": DeviceArray((m, m)),
             "Cvv": DeviceArray((nact, nact)), "Tmp": DeviceArray((nact, m))}
        check(lib().moao_hip_cee_cvv_dev(
            n, m, nact, vp(dCmm.ptr), n, vp(s["Cpp"].ptr), m, vp(s["Cpm"].ptr), m,
            vp(s["R"].ptr), m, vp(dDx.ptr), nact, vp(s["Cee"].ptr), m, vp(s["Cvv"].ptr), nact,
            vp(s["Tmp"].ptr), nact, None), "cee_cvv")
        for k, v in s.items():
            w = width[k]
            assert np.array_equal(got[k][:, g * w:(g + 1) * w], v.to_host()), (g, k)
    # step 1 (dsyr2k) is one product X = Cpm R^T and Cpp -= X, Cpp -= X^T: its lower part
    # equals the two accumulating products Cpp -= Cpm R^T, Cpp -= R Cpm^T bit for bit
    from moao_amd.reconstructor import dgemm_dev
    for g in (0, ngal - 1):
        dC, dP, dRg = DeviceArray.from_host(Ctt[g]), DeviceArray.from_host(Ctm[g]), \
            DeviceArray.from_host(R[g])
        dgemm_dev(1, m, m, n, -1.0, dP.ptr, m, dRg.ptr, m, 1.0, dC.ptr, m)
        dgemm_dev(1, m, m, n, -1.0, dRg.ptr, m, dP.ptr, m, 1.0, dC.ptr, m)
        low = np.tril_indices(m)
        assert np.array_equal(got["Cpp"][:, g * m:(g + 1) * m][low], dC.to_host()[low]), g
        up = np.triu_indices(m, 1)
        assert np.array_equal(got["Cpp"][:, g * m:(g + 1) * m][up], Ctt[g][up]), g
    # and one target against the closed form (the per-target path's own parity is above)
    Cl = np.tril(Cmm) + np.tril(Cmm, -1).T
    g = ngal - 1
    Cee_cf = Ctt[g] - Ctm[g] @ R[g].T - R[g] @ Ctm[g].T + R[g] @ Cl @ R[g].T
    Cee_g = got["Cee"][:, g * m:(g + 1) * m]
    assert rel_fro(Cee_g, Cee_cf) <= 1e-12
