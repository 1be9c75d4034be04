"""Serving engine: continuous batching of reconstructor requests of uneven size.

Bars:
  * each request's R is bit-identical whether it runs alone (one slot, one request at a time)
    or in a mix with others overlapping on the device (same kernels, same order, same inputs);
  * FP64 requests match the oracle's LAPACK path on the same Cmm/Ctm within 1e-10 relative
    Frobenius error (the FP64 parity bar); mixed requests within 1e-6 of it;
  * admission never lets the requests in flight hold more than the byte budget, and a request
    that alone exceeds the budget is refused at submit with MOAO_ERR_NOMEM;
  * a request that fails on the device (a mixed solve that cannot converge) reports its status
    and leaves the requests beside it, and the ones after it, exactly as they are alone.
"""
import numpy as np
import pytest

from moao_amd import matcov as gm
from moao_amd import systems
from moao_amd.lib import (MOAO_ERR_NOCONV, MOAO_ERR_NOMEM, MOAO_FP64, MOAO_MIXED_FP32,
                          MoaoError)
from moao_amd.serve import Server
from moao_amd.tomo import Tomo
from oracle import lapack_path as lp
from oracle.check import rel_fro

pytestmark = pytest.mark.gpu

# uneven systems: (nssp, nlgs, nngs, nlayer, diam, seed)
SPECS = [(12, 3, 0, 2, 8.0, 11), (20, 4, 1, 4, 8.0, 12), (16, 2, 1, 3, 8.0, 13),
         (24, 3, 1, 4, 10.0, 14), (10, 3, 0, 2, 8.0, 15), (18, 4, 0, 3, 8.0, 16)]


def _systems(tmp_path):
    out = []
    for k, (nssp, nlgs, nngs, nl, diam, seed) in enumerate(SPECS):
        spec = systems.SystemSpec(diam=diam, nssp=nssp, nlgs=nlgs, nngs=nngs,
                                  lgs_ring_diam_as=60.0, ngs_field_diam_as=60.0, nlayer=nl,
                                  r0=0.15, seed=seed)
        d = tmp_path / f"sys{k}"
        d.mkdir()
        s, atms = systems.make_system(str(d), spec)
        t = Tomo(s, atms[0])
        out.append((s, atms[0], t.nmeas(), t.nmeasts()))
    return out


def _mode(k):
    return MOAO_FP64 if k % 2 == 0 else MOAO_MIXED_FP32


def test_mix_equals_alone_and_oracle(tmp_path):
    syst = _systems(tmp_path)
    print("[serve] sizes", [(n, m) for _, _, n, m in syst], flush=True)
    with Server(slots=3) as srv:
        tickets = [srv.submit(s, a, 0, _mode(k), n_m=(n, m))
                   for k, (s, a, n, m) in enumerate(syst)]
        mix = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print(f"[serve] mix: peak_running={st['peak_running']} peak_bytes={st['peak_bytes']}",
          flush=True)
    assert st["peak_running"] >= 2  # requests did overlap on the device
    alone = []
    with Server(slots=1) as srv:
        for k, (s, a, n, m) in enumerate(syst):
            alone.append(srv.wait(srv.submit(s, a, 0, _mode(k), n_m=(n, m))))
    for k, (s, a, n, m) in enumerate(syst):
        assert mix[k]["status"] == 0 and alone[k]["status"] == 0
        assert np.array_equal(mix[k]["R"], alone[k]["R"]), f"request {k} differs in the mix"
        # the oracle's LAPACK path on the matrices the library's matcov generates (the same
        # kernel the server runs), so only the solve is compared
        t = Tomo(s, a)
        C = gm.cmm(t)
        t.set_target(0)
        R = gm.ctm(t).copy(order="F")
        assert lp.reconstructor(C.copy(order="F"), R) == 0
        err = rel_fro(mix[k]["R"], R)
        tol = 1e-10 if _mode(k) == MOAO_FP64 else 1e-6
        print(f"    request {k} n={n} m={m} mode={_mode(k)} err={err:.2e} "
              f"latency={mix[k]['latency_s'] * 1e3:.1f} ms slot={mix[k]['slot']}", flush=True)
        assert err <= tol


def test_admission_respects_budget(tmp_path):
    syst = _systems(tmp_path)
    with Server(slots=4) as probe:
        fp = [probe.wait(probe.submit(s, a, 0, MOAO_FP64, want_R=False))["footprint"]
              for s, a, _, _ in syst]
    budget = int(max(fp) * 1.3)
    print(f"[serve] budget {budget} B, footprints {fp}", flush=True)
    with Server(slots=4, hbm_budget=budget) as srv:
        tickets = [srv.submit(s, a, 0, MOAO_FP64, want_R=False) for s, a, _, _ in syst]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
        assert all(r["status"] == 0 for r in res)
        assert st["peak_bytes"] <= budget
        print(f"[serve] peak_bytes={st['peak_bytes']} peak_running={st['peak_running']}",
              flush=True)
    with Server(slots=1, hbm_budget=min(fp) // 2) as srv:
        s, a, _, _ = syst[0]
        with pytest.raises(MoaoError) as e:
            srv.submit(s, a, 0, MOAO_FP64, want_R=False)
        assert e.value.code == MOAO_ERR_NOMEM


def test_failed_request_is_isolated(tmp_path):
    syst = _systems(tmp_path)[:4]
    print("[serve] failing request in a mix", flush=True)
    with Server(slots=3) as srv:
        tickets = []
        for k, (s, a, n, m) in enumerate(syst):
            if k == 1:  # cannot reach the tolerance in one refinement step
                tickets.append(srv.submit(s, a, 0, MOAO_MIXED_FP32, tol=1e-30, max_iter=1,
                                          n_m=(n, m)))
            else:
                tickets.append(srv.submit(s, a, 0, MOAO_FP64, n_m=(n, m)))
        res = [srv.wait(t, raise_on_error=False) for t in tickets]
        s, a, n, m = syst[1]  # the server keeps serving after the failure
        after = srv.wait(srv.submit(s, a, 0, MOAO_FP64, n_m=(n, m)), raise_on_error=False)
        st = srv.stats()
    assert res[1]["status"] == MOAO_ERR_NOCONV and res[1]["R"] is None
    assert after["status"] == 0
    assert st["peak_running"] >= 2
    with Server(slots=1) as srv:
        alone = {k: srv.wait(srv.submit(s, a, 0, MOAO_FP64, n_m=(n, m)))
                 for k, (s, a, n, m) in enumerate(syst) if k != 1}
    for k, r in alone.items():
        assert res[k]["status"] == 0
        assert np.array_equal(res[k]["R"], r["R"]), f"request {k} differs beside a failure"
    s, a, n, m = syst[1]
    with Server(slots=1) as srv:
        assert np.array_equal(after["R"], srv.wait(srv.submit(s, a, 0, MOAO_FP64, n_m=(n, m)))["R"])


def test_repeated_shapes_replay_graphs(tmp_path):
    """A stream of refine steps of one system (same shape, fresh atmospheric profiles) on one
    slot: the first FP64 solve runs eagerly, the second is recorded as a graph (the pool handed
    back the same block) and the later ones replay it.  Every R equals the one computed by the
    eager entry on the same matcov inputs, bit for bit."""
    from moao_amd.device import DeviceArray, synchronize
    from moao_amd.reconstructor import reconstructor_dev
    spec = systems.SystemSpec(diam=8.0, nssp=20, nlgs=4, nngs=1, lgs_ring_diam_as=60.0,
                              ngs_field_diam_as=60.0, nlayer=4, r0=0.15, seed=21)
    s, atms = systems.make_system(str(tmp_path), spec, nprof=5)
    t0 = Tomo(s, atms[0])
    n, m = t0.nmeas(), t0.nmeasts()
    print(f"[serve-graph] n={n} m={m}, {len(atms)} profiles", flush=True)
    with Server(slots=1) as srv:
        res = [srv.wait(srv.submit(s, a, 0, MOAO_FP64, n_m=(n, m))) for a in atms]
        st = srv.stats()
    print(f"    captures={st['graph_captures']} replays={st['graph_replays']}", flush=True)
    assert st["graph_captures"] + st["graph_replays"] <= len(atms) - 1
    assert st["graph_captures"] >= 1 and st["graph_replays"] >= 1
    for a, r in zip(atms, res):
        t = Tomo(s, a)
        t.set_target(0)
        dC, dB = DeviceArray.from_host(gm.cmm(t)), DeviceArray.from_host(gm.ctm(t))
        assert reconstructor_dev(n, m, dC.ptr, n, dB.ptr, m, 1) == 0
        synchronize()
        assert np.array_equal(r["R"], dB.to_host())
