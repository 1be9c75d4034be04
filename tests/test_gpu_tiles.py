"""Tile descriptors and sequences (the MOAO_CHAMELEON::reconstructor call surface,
chameleon/compute/reconstructor.cpp:31-65): Cmm and Ctm converted to tile layout with tile size
ts, the asynchronous solve queued on a sequence, waited on, converted back.  The solve runs the
same kernels in the same order as the column-major entry on the same data, so the factor and R
must be bit-identical to moao_hip_reconstructor_dev (FP64) and moao_hip_sreconstructor_dev (FP32
storage, the Chameleon production precision); a non-SPD Cmm reports its pivot through the
sequence's status."""
import ctypes

import numpy as np
import pytest

from moao_amd.device import DeviceArray, synchronize
from moao_amd.lib import TileDesc, check, lib

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def _system(n, m, seed, dtype, spd=True):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n)) / np.sqrt(n)
    C = G @ G.T + np.eye(n)
    if not spd:
        C[n // 2, n // 2] = -1.0
    B = rng.standard_normal((m, n))
    return np.asfortranarray(np.tril(C).astype(dtype)), np.asfortranarray(B.astype(dtype))


def _desc(rows, cols, ts, dtype):
    mt, nt = -(-rows // ts), -(-cols // ts)
    buf = DeviceArray((mt * nt * ts * ts,), dtype)
    return TileDesc(0 if dtype == np.float64 else 1, buf.ptr, ts, ts, rows, cols), buf


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ts,sync", [(200, 0), (128, 1), (512, 0)])
def test_tile_solve_matches_column_major_entry(dtype, ts, sync):
    n, m = 1000, 217
    print(f"[tiles] {np.dtype(dtype).name} n={n} m={m} ts={ts} sync={sync}", flush=True)
    C, B = _system(n, m, 3, dtype)
    # reference: the column-major entry
    dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
    f = (lib().moao_hip_reconstructor_dev if dtype == np.float64
         else lib().moao_hip_sreconstructor_dev)
    assert f(n, m, vp(dC.ptr), n, vp(dB.ptr), m, 1, None) == 0
    L_ref, R_ref = dC.to_host(), dB.to_host()
    # tile path
    dC2, dB2 = DeviceArray.from_host(C), DeviceArray.from_host(B)
    q = ctypes.c_void_p()
    check(lib().moao_hip_sequence_create(ctypes.byref(q)), "sequence_create")
    try:
        dcm, bufc = _desc(n, n, ts, dtype)
        dct, bufb = _desc(m, n, ts, dtype)
        check(lib().moao_hip_lapack_to_tile(vp(dC2.ptr), n, ctypes.byref(dcm), q), "to_tile")
        check(lib().moao_hip_lapack_to_tile(vp(dB2.ptr), m, ctypes.byref(dct), q), "to_tile")
        check(lib().moao_hip_reconstructor_tile_async(ctypes.byref(dcm), ctypes.byref(dct), q,
                                                      0, sync), "reconstructor_tile_async")
        out_c, out_b = DeviceArray((n, n), dtype), DeviceArray((m, n), dtype)
        check(lib().moao_hip_tile_to_lapack(ctypes.byref(dcm), vp(out_c.ptr), n, q), "to_lapack")
        check(lib().moao_hip_tile_to_lapack(ctypes.byref(dct), vp(out_b.ptr), m, q), "to_lapack")
        assert lib().moao_hip_sequence_wait(q) == 0
        assert np.array_equal(out_b.to_host(), R_ref)
        assert np.array_equal(np.tril(out_c.to_host()), np.tril(L_ref))
    finally:
        lib().moao_hip_sequence_destroy(q)


def test_tile_solve_reports_pivot_through_the_sequence():
    n, m, ts = 600, 50, 100
    C, B = _system(n, m, 4, np.float64, spd=False)
    dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
    q = ctypes.c_void_p()
    check(lib().moao_hip_sequence_create(ctypes.byref(q)), "sequence_create")
    try:
        dcm, bufc = _desc(n, n, ts, np.float64)
        dct, bufb = _desc(m, n, ts, np.float64)
        check(lib().moao_hip_lapack_to_tile(vp(dC.ptr), n, ctypes.byref(dcm), q), "to_tile")
        check(lib().moao_hip_lapack_to_tile(vp(dB.ptr), m, ctypes.byref(dct), q), "to_tile")
        check(lib().moao_hip_reconstructor_tile_async(ctypes.byref(dcm), ctypes.byref(dct), q,
                                                      0, 0), "reconstructor_tile_async")
        st = lib().moao_hip_sequence_wait(q)
        assert 0 < st <= n // 2 + 1
    finally:
        lib().moao_hip_sequence_destroy(q)
        synchronize()
