"""Mixed-precision reconstructor (low-precision factorization of C0 = Cmm - lgs_cst U U^T,
Woodbury in FP64, FP64 refinement against the full Cmm) against the FP64 lapack-path oracle, in
both modes: FP32 MFMA updates, and split-BF16 updates (each FP32 operand as two BF16 terms on
BF16 MFMA, FP32 accumulation).

Bar: relative Frobenius error of R <= 1e-6 (the accuracy the mixed path is specified to), with
the refinement stopping at ||Ctm - R Cmm|| / ||Ctm|| <= 1e-11; Cmm must be left untouched.
"""
import numpy as np
import pytest

from conftest import FIXTURES, fixture_system
from moao_amd import matcov as gm
from moao_amd import systems
from moao_amd.device import DeviceArray
from moao_amd.lib import MOAO_MIXED_BF16, MOAO_MIXED_FP32
from moao_amd.reconstructor import reconstructor_mixed_dev
from moao_amd.tomo import Tomo
from oracle import lapack_path as lp
from oracle.check import rel_fro

pytestmark = pytest.mark.gpu
MODES = [MOAO_MIXED_FP32, MOAO_MIXED_BF16]


def _run(t, use_lowrank=True, tol=1e-11, max_iter=6, mode=MOAO_MIXED_FP32):
    C = gm.cmm(t)
    t.set_target(0)
    B = gm.ctm(t)
    n, m = C.shape[0], B.shape[0]
    Ro = B.copy(order="F")
    assert lp.reconstructor(C.copy(order="F"), Ro) == 0
    dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
    rc, res, its = reconstructor_mixed_dev(n, m, dC.ptr, n, dB.ptr, m,
                                           t if use_lowrank else None, tol, max_iter, mode=mode)
    R = dB.to_host()
    assert np.array_equal(dC.to_host(), C)  # Cmm preserved
    return rel_fro(R, Ro), res, its


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", FIXTURES)
def test_mixed_fixture_systems(name, mode):
    print(f"[mixed] fixture {name} mode {mode}", flush=True)
    err, res, its = _run(Tomo(*fixture_system(name)), mode=mode)
    print(f"    err={err:.2e} resid={res:.2e} steps={its}", flush=True)
    assert res <= 1e-11 and err <= 1e-6


@pytest.mark.parametrize("mode", MODES)
def test_mixed_small_workload(tmp_path, mode):
    print(f"[mixed] seeded small workload mode {mode}", flush=True)
    s, atms = systems.make_system(str(tmp_path), "small")
    err, res, its = _run(Tomo(s, atms[0]), mode=mode)
    print(f"    err={err:.2e} resid={res:.2e} steps={its}", flush=True)
    assert res <= 1e-11 and err <= 1e-6


@pytest.fixture(scope="module")
def elt_system(tmp_path_factory):
    """n = 18,688 (the 40x40, 6 LGS + 2 NGS, 10-layer system): Cmm, Ctm from the library's
    matcov and the oracle's R on them."""
    d = tmp_path_factory.mktemp("elt")
    s, atms = systems.make_system(str(d), "elt")
    t = Tomo(s, atms[0])
    C = gm.cmm(t)
    t.set_target(0)
    B = gm.ctm(t)
    Ro = B.copy(order="F")
    assert lp.reconstructor(C.copy(order="F"), Ro) == 0
    return t, C, B, Ro


def test_fp64_elt_workload(elt_system):
    """The plain FP64 GPU solve of the n = 18,688 system against the oracle."""
    from moao_amd.reconstructor import reconstructor
    print("[fp64] seeded elt workload (n=18688)", flush=True)
    t, C, B, Ro = elt_system
    n, m = C.shape[0], B.shape[0]
    Cg, Bg = C.copy(order="F"), B.copy(order="F")
    assert reconstructor(n, m, Cg, n, Bg, m, 1) == 0
    err = rel_fro(Bg, Ro)
    print(f"    err={err:.2e}", flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("mode", MODES)
def test_mixed_elt_workload(elt_system, mode):
    """n = 18,688: the factorization and solves run on the 128-row FP32 (or split-BF16 planes)
    kernels and every FP64 residual on the symmetric direct-to-LDS kernel."""
    print(f"[mixed] seeded elt workload (n=18688) mode {mode}", flush=True)
    t, C, B, Ro = elt_system
    n, m = C.shape[0], B.shape[0]
    dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
    rc, res, its = reconstructor_mixed_dev(n, m, dC.ptr, n, dB.ptr, m, t, 1e-11, 6, mode=mode)
    R = dB.to_host()
    assert np.array_equal(dC.to_host(), C)
    err = rel_fro(R, Ro)
    print(f"    err={err:.2e} resid={res:.2e} steps={its}", flush=True)
    assert res <= 1e-11 and err <= 1e-6


def test_mixed_without_lowrank_split(tmp_path):
    """No low-rank description: FP32 factorization of Cmm itself (small n, still SPD in FP32)."""
    print("[mixed] no low-rank split", flush=True)
    s, atms = systems.make_system(str(tmp_path), "small")
    err, res, its = _run(Tomo(s, atms[0]), use_lowrank=False, max_iter=10)
    print(f"    err={err:.2e} resid={res:.2e} steps={its}", flush=True)
    assert res <= 1e-11 and err <= 1e-6


def test_mixed_reports_nonconvergence(tmp_path):
    from moao_amd.lib import MOAO_ERR_NOCONV
    print("[mixed] max_iter=0 reports non-convergence", flush=True)
    s, atms = systems.make_system(str(tmp_path), "small")
    t = Tomo(s, atms[0])
    C = gm.cmm(t)
    t.set_target(0)
    B = gm.ctm(t)
    dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
    rc, res, its = reconstructor_mixed_dev(C.shape[0], B.shape[0], dC.ptr, C.shape[0], dB.ptr,
                                           B.shape[0], t, 1e-15, 0, allow_noconv=True)
    assert rc == MOAO_ERR_NOCONV and its == 0 and res > 1e-15
