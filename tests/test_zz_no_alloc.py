"""The allocation-free reconstructor entry (moao_hip_reconstructor_noalloc: its workspace is a
static device array) against the stream-ordered entry that allocates its workspace
(moao_hip_reconstructor_async), through the C-ABI.

Both run the same kernels in the same order on the same inputs, so R, the factor and the status
must be bit-identical.  Two calls in a row share the scratch: the second, on a smaller system,
must not see anything the first left there.  A non-SPD input must report the same pivot, and an
order beyond the scratch maximum is refused before any device work."""
import ctypes

import numpy as np
import pytest

from moao_amd.device import DeviceArray, synchronize
from moao_amd.lib import MOAO_ERR_ARG, lib

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def _system(n, m, seed, spd=True):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((n, n)) / np.sqrt(n)
    C = G @ G.T + np.diag(rng.uniform(0.5, 1.5, n))
    if not spd:
        C[n // 2, n // 2] = -1.0
    B = rng.standard_normal((m, n))
    return np.asfortranarray(C), np.asfortranarray(B)


def _solve(entry, C, B):
    n, m = C.shape[0], B.shape[0]
    dC, dB = DeviceArray.from_host(C), DeviceArray.from_host(B)
    dinfo = DeviceArray((1,), np.int32)
    rc = getattr(lib(), entry)(n, m, vp(dC.ptr), n, vp(dB.ptr), m, 1, vp(dinfo.ptr), None)
    assert rc == 0, (entry, rc, lib().moao_last_error())
    synchronize()
    return dC.to_host(), dB.to_host(), int(dinfo.to_host()[0])


@pytest.mark.parametrize("spd", [True, False])
def test_noalloc_matches_allocating_entry_on_two_calls(spd):
    # second call smaller than the first, with a ragged last leaf both times
    for k, (n, m) in enumerate([(3000 + 37, 517), (700 + 5, 131)]):
        print(f"[noalloc] call {k}: n={n} m={m} spd={spd}", flush=True)
        C, B = _system(n, m, seed=11 + k, spd=spd)
        L_ref, R_ref, info_ref = _solve("moao_hip_reconstructor_async", C, B)
        L, R, info = _solve("moao_hip_reconstructor_noalloc", C, B)
        assert info == info_ref
        assert (info == 0) == spd
        assert np.array_equal(np.tril(L), np.tril(L_ref))
        assert np.array_equal(R, R_ref)
        if spd:
            S = np.tril(C) + np.tril(C, -1).T
            bwd = np.linalg.norm(B - R @ S) / np.linalg.norm(B)
            assert bwd <= 5e-13


def test_noalloc_refuses_orders_beyond_its_scratch():
    d = DeviceArray((64,))
    dinfo = DeviceArray((1,), np.int32)
    n = 40961
    rc = lib().moao_hip_reconstructor_noalloc(n, 1, vp(d.ptr), n, vp(d.ptr), 1, 1, vp(dinfo.ptr),
                                               None)
    assert rc == MOAO_ERR_ARG
    assert b"40960" in lib().moao_last_error()
