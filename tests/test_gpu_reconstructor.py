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
    """ldCmm > n on a buffer of exactly LAPACK's minimum extent ((n-1)*ldCmm + n doubles), and
    two targets stored n*m apart (reconstructor.cpp:43)."""
    print("[recon] ldCmm > n (minimum buffer) + 2 targets", flush=True)
    rng = np.random.default_rng(5)
    n, m, ld_a = 300, 40, 310
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    Cbuf = np.zeros((n - 1) * ld_a + n)
    for c in range(n):
        Cbuf[c * ld_a + c:c * ld_a + n] = S[c:, c]
    Cbuf[[c * ld_a + r for c in range(n - 1) for r in range(n, ld_a)]] = 7.0  # row padding
    B0, B1 = rng.standard_normal((m, n)), rng.standard_normal((m, n))
    Bbuf = np.concatenate([np.asfortranarray(B0).reshape(-1, order="F"),
                           np.asfortranarray(B1).reshape(-1, order="F")])
    info = reconstructor(n, m, Cbuf, ld_a, Bbuf, m, 2)
    assert info == 0
    for g, Bg in enumerate((B0, B1)):
        R = Bbuf[g * n * m:(g + 1) * n * m].reshape((m, n), order="F")
        assert np.linalg.norm(Bg - R @ S) / np.linalg.norm(Bg) <= 5e-13
    # padding rows between the columns come back unchanged
    assert np.all(Cbuf[[c * ld_a + r for c in range(n - 1) for r in range(n, ld_a)]] == 7.0)


def test_ldctm_greater_than_m():
    """One target with ldCtm > m on a buffer of exactly (n-1)*ldCtm + m doubles."""
    print("[recon] ldCtm > m (minimum buffer)", flush=True)
    rng = np.random.default_rng(6)
    n, m, ld_b = 260, 30, 37
    G = rng.standard_normal((n, n))
    S = G @ G.T / n + np.eye(n)
    C = np.asfortranarray(np.tril(S))
    B = rng.standard_normal((m, n))
    Bbuf = np.full((n - 1) * ld_b + m, 5.0)
    for c in range(n):
        Bbuf[c * ld_b:c * ld_b + m] = B[:, c]
    assert reconstructor(n, m, C, n, Bbuf, ld_b, 1) == 0
    R = np.stack([Bbuf[c * ld_b:c * ld_b + m] for c in range(n)], axis=1)
    assert np.linalg.norm(B - R @ S) / np.linalg.norm(B) <= 5e-13
    assert np.all(Bbuf[[c * ld_b + r for c in range(n - 1) for r in range(m, ld_b)]] == 5.0)


def test_overlapping_targets_are_rejected():
    """With ngal > 1 the blocks start g*n*m apart; ldCtm > m would make them overlap."""
    from moao_amd.lib import MOAO_ERR_ARG, MoaoError
    n, m = 64, 8
    C = np.asfortranarray(np.eye(n))
    B = np.zeros(2 * n * 10)
    with pytest.raises(MoaoError) as e:
        reconstructor(n, m, C, n, B, 10, 2)
    assert e.value.code == MOAO_ERR_ARG


def _partial_factor_ok(Cg, S, ncols, leaf0):
    """After a failing pivot in the leaf starting at column leaf0, the factor is complete in
    every column of the earlier leaves (all rows) and, in the failing leaf's 16-column blocks
    before the failing block (columns leaf0..ncols-1), in the leaf's own rows: the panel rows
    below a leaf are solved only once the whole leaf is factored.  Reference columns: the
    Cholesky factor of the leading block (numpy) and the rows below by a triangular solve."""
    L11 = np.linalg.cholesky(S[:ncols, :ncols])
    L21 = np.linalg.solve(L11, S[ncols:, :ncols].T).T
    ref = np.vstack([L11, L21])
    got = np.tril(Cg[:, :ncols])
    full = slice(0, leaf0)
    assert np.linalg.norm(got[:, full] - ref[:, full]) / np.linalg.norm(ref[:, full]) <= 1e-12
    rows = slice(leaf0, min(leaf0 + 128, S.shape[0]))
    part = slice(leaf0, ncols)
    assert (np.linalg.norm(got[rows, part] - ref[rows, part]) /
            np.linalg.norm(ref[rows, part]) <= 1e-12)


def test_two_streams_overlap():
    """Two FP64 solves queued on two streams with the stream-ordered entry (their kernels may
    overlap on the device): each R is bit-identical to the same solve run alone."""
    from moao_amd.device import DeviceArray, Stream, synchronize
    from moao_amd.reconstructor import reconstructor_async, reconstructor_dev
    print("[recon] two streams", flush=True)
    rng = np.random.default_rng(21)
    probs = []
    for n, m in [(3000, 400), (2200, 700)]:
        G = rng.standard_normal((n, n))
        C = np.asfortranarray(np.tril(G @ G.T / n + np.eye(n)))
        B = np.asfortranarray(rng.standard_normal((m, n)))
        probs.append((n, m, C, B))
    solo = []
    for n, m, C, B in probs:
        dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
        assert reconstructor_dev(n, m, dC.ptr, n, dB.ptr, m, 1) == 0
        solo.append(dB.to_host())
    s1, s2 = Stream(), Stream()
    dev = [(DeviceArray.from_host(C), DeviceArray.from_host(B)) for _, _, C, B in probs]
    info = DeviceArray((2,), np.int32)
    synchronize()
    for k, ((n, m, _, _), (dC, dB), st) in enumerate(zip(probs, dev, (s1, s2))):
        reconstructor_async(n, m, dC.ptr, n, dB.ptr, m, 1, info.ptr + 4 * k, st.ptr)
    s1.synchronize()
    s2.synchronize()
    assert list(info.to_host()) == [0, 0]
    for (dC, dB), R in zip(dev, solo):
        assert np.array_equal(dB.to_host(), R)
    s1.destroy()
    s2.destroy()


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
    Cg = C.copy(order="F")
    info = reconstructor(n, m, Cg, n, Bg, m, 1)
    assert info == info_ref == 301
    assert np.array_equal(Bg, B)  # R not computed: Ctm untouched on the host entry
    # column 300 is local column 44 of the leaf [256, 384): the leaves before it and, in that
    # leaf's rows, its 16-column blocks before column 288 hold the factor
    S = np.tril(C) + np.tril(C, -1).T
    _partial_factor_ok(Cg, S, 288, 256)


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
    Cg = C.copy(order="F")
    assert reconstructor(n, m, Cg, n, Bg, m, 1) == 201
    assert np.array_equal(Bg, B)
    # leaf [128, 256): its blocks before column 192 (the one holding column 200) are stored in
    # the leaf's rows.  The NaN sits in the block [144, 160), whose rows are formed as one
    # product with the block's inverse (0 * NaN = NaN), so that block's row 200 is NaN in
    # every column; the blocks before it are checked.
    S = np.tril(C) + np.tril(C, -1).T
    _partial_factor_ok(Cg, S, 144, 128)


def test_solve_graph_replays_match_eager():
    """The FP64 solve recorded as a graph (moao_hip_solve_graph_*) and replayed three times on
    refilled buffers -- two different systems of the same shape, eager solves and copies of the
    program's own in between -- gives R and the factor bit-identical to the eager solve each
    time."""
    import ctypes
    from moao_amd.device import DeviceArray, synchronize
    from moao_amd.lib import check, lib
    from moao_amd.reconstructor import reconstructor_dev
    print("[recon] solve graph", flush=True)
    rng = np.random.default_rng(31)
    n, m = 2100, 333
    probs = []
    for _ in range(2):
        G = rng.standard_normal((n, n))
        C = np.asfortranarray(np.tril(G @ G.T / n + np.eye(n)))
        probs.append((C, np.asfortranarray(rng.standard_normal((m, n)))))
    eager = []
    for C, B in probs:
        dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
        assert reconstructor_dev(n, m, dC.ptr, n, dB.ptr, m, 1) == 0
        eager.append((dC.to_host(), dB.to_host()))
    src = [(DeviceArray.from_host(C), DeviceArray.from_host(B)) for C, B in probs]
    dC, dB = DeviceArray((n, n)), DeviceArray((m, n))
    info = DeviceArray((1,), np.int32)
    g = ctypes.c_void_p()
    vp = ctypes.c_void_p
    check(lib().moao_hip_solve_graph_create(n, m, vp(dC.ptr), n, vp(dB.ptr), m, 1,
                                            vp(info.ptr), ctypes.byref(g)),
          "moao_hip_solve_graph_create")
    try:
        for rep, k in enumerate([0, 1, 0]):
            dC.copy_from(src[k][0])
            dB.copy_from(src[k][1])
            check(lib().moao_hip_solve_graph_launch(g, None), "moao_hip_solve_graph_launch")
            synchronize()
            assert int(info.to_host()[0]) == 0
            assert np.array_equal(np.tril(dC.to_host()), np.tril(eager[k][0])), rep
            assert np.array_equal(dB.to_host(), eager[k][1]), rep
            # eager work of the program between replays
            e1, e2 = DeviceArray.from_host(probs[1 - k][0]), DeviceArray.from_host(probs[1 - k][1])
            assert reconstructor_dev(n, m, e1.ptr, n, e2.ptr, m, 1) == 0
    finally:
        lib().moao_hip_solve_graph_destroy(g)
