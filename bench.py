"""Benchmark of the MOAO reconstructor hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload elt74]
                    [--modes fp64,mixed,bf16] [--serve 256] [--serve-modes fp64]
                    [--serve-mixed 128] [--obs-loop 1] [--obs-prepare DIR]
                    [--slots 4] [--no-cpu]

One step = one R-solve of the workload: restore the resident inputs (device-to-device copies of
the Cmm and Ctm generated once by matcov) and solve R = Ctm * Cmm^{-1} in place.  The inputs
are in HBM before the timed region.  Prints ONE JSON line (rank 0).

value           R-solves/s of the headline path: the fastest of FP64 and the mixed modes
                (mixed = FP32 MFMA updates, bf16 = split-BF16 updates, both refined against
                FP64 residuals) whose R is within 1e-6 of the FP64 R, timed over exactly K
                steps after W warm-up steps
latency         p50 / p99 of the per-step wall time (each step synchronised)
roofline        FP64 MFMA GEMM engine: algorithmic GEMM flops / summed GEMM kernel time
                (HIP events on the launch stream) vs the 78.6 TFLOP/s FP64 peak
serving         a Poisson stream of uneven systems (n up to 40k) through the serving engine:
                throughput and p50/p99 latency from submit to R on the host, FP64 (256
                requests; split-BF16 mixed, 128, with --serve-modes fp64,bf16), with the slots'
                graph captures and replays
observation_loop  the refine/observation loop of the `moao` driver on the n = 18,688 system with
                8 directions (4 refine x 16 observation steps, Cee/Cvv, batched PSFs, long
                exposure), and the PSF Strehl with FP32 and FP64 transforms
cpu_baseline    the oracle's LAPACK path (scipy dpotrf + 2 dtrsm, OpenBLAS) on the same full
                matrices, all host cores; the oracle's matcov on a bounded column sample, one
                thread and all host threads; the oracle's FP64 PSF of the loop's last step
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "R-solves/s + p50/p99 latency, mixed vs FP64 at <=1e-6 rel err, n=40k"
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12
PEAK_BF16 = 2.5e15


def solve_flops(n, m):
    """LAWN-41 POTRF + 2 right-side TRSM (chameleon/flops.h:72,80,364)."""
    potrf = n * ((n / 6.0 + 0.5) * n + 1.0 / 3.0) + n * ((n / 6.0) * n - 1.0 / 6.0)
    trsm = 2.0 * 0.5 * m * n * (n + 1.0)
    return potrf + 2.0 * trsm


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([p.get("num_threads", 1) for p in threadpool_info()
                    if p.get("internal_api") in ("openblas", "mkl", "blis")] or [1])
    except Exception:
        return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(dCmm0, dCtm0, n, m):
    """The oracle's LAPACK path (scipy dpotrf + 2 dtrsm in the reference's call order) on the
    full Cmm and Ctm of the workload, all host BLAS threads: one timed R-solve."""
    from moao_amd.lib import lib
    import ctypes
    from oracle import lapack_path
    threads = _blas_threads()
    C = np.empty((n, n), order="F")
    lib().moao_hip_memcpy_d2h(C.ctypes.data, ctypes.c_void_p(dCmm0.ptr), C.nbytes)
    B = np.empty((m, n), order="F")
    lib().moao_hip_memcpy_d2h(B.ctypes.data, ctypes.c_void_p(dCtm0.ptr), B.nbytes)
    t0 = time.perf_counter()
    info = lapack_path.reconstructor(C, B)
    dt = time.perf_counter() - t0
    del C, B
    return {"value": 1.0 / dt, "unit": "R-solves/s", "cores": threads, "kind": "port",
            "sample": f"one full n={n}, m={m} R-solve: scipy/OpenBLAS dpotrf + 2 dtrsm on the "
                      f"workload's own Cmm and Ctm ({dt:.1f} s, "
                      f"{solve_flops(n, m) / dt / 1e9:.0f} GFLOP/s, info={info})"}


def cpu_matcov(sys_file, atm_file, n, m, cols=48, threads=None):
    """The oracle's matcov (the reference's element function restated in C) on a column sample
    of Cmm at the middle of the matrix: one thread, then all host threads each on its own
    column block (ctypes releases the GIL); elements/s and the time the whole Cmm-lower + Ctm
    would take at that rate."""
    import concurrent.futures as cf
    from oracle import matcov as om
    from oracle.tomo import Tomo as OTomo
    o = OTomo.from_files(sys_file, atm_file)
    o.set_target(0)
    threads = threads or min(16, os.cpu_count() or 1)
    c0 = n // 2
    rows = n - c0

    def block(k):
        buf = np.zeros((rows, cols), order="F")
        om.matcov_tile(buf, rows, cols, c0, c0 + k * cols, rows, o, 1)
        return sum(max(0, n - c) for c in range(c0 + k * cols, c0 + (k + 1) * cols))

    t0 = time.perf_counter()
    e1 = block(0)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        eN = sum(ex.map(block, range(threads)))
    tN = time.perf_counter() - t0
    total = n * (n + 1) / 2 + m * n
    return {"elements_total": total,
            "serial": {"elements_per_s": e1 / t1, "full_s": total / (e1 / t1), "threads": 1,
                       "sample": f"{cols} Cmm columns from {c0} ({e1} elements, {t1:.1f} s)"},
            "threaded": {"elements_per_s": eN / tN, "full_s": total / (eN / tN),
                         "threads": threads,
                         "sample": f"{threads} x {cols} Cmm columns from {c0} ({eN} elements, "
                                   f"{tN:.1f} s)"}}


def hbm_traffic(workload):
    """HBM bytes per FP64 GEMM launch from the two calibrated rocprofv3 counter passes
    (FETCH_SIZE, WRITE_SIZE) recorded for this workload, or None if none were recorded."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    return rec["bytes_per_launch"] if rec else None


def cee_of(tomo, n, m, dCmm0, dCtm0, R, nact=64):
    """Cee of the reconstructor R (host array) on the workload's Cmm, Ctm and Ctt, through the
    Cee/Cvv entry (a small Dx: only Cee is used)."""
    import ctypes
    from moao_amd import systems
    from moao_amd.device import DeviceArray, synchronize
    from moao_amd.lib import check, lib
    from moao_amd.matcov import CTT, matcov_tile_dev
    vp = ctypes.c_void_p
    dCpp, dCpm = DeviceArray((m, m)), DeviceArray((m, n))
    matcov_tile_dev(dCpp.ptr, m, m, 0, 0, m, tomo, CTT)
    dCpm.copy_from(dCtm0)
    dR = DeviceArray.from_host(R)
    dDx = DeviceArray.from_host(systems.dx_matrix(nact, m, n))
    dCee, dCvv, dTmp = DeviceArray((m, m)), DeviceArray((nact, nact)), DeviceArray((nact, m))
    check(lib().moao_hip_cee_cvv_dev(n, m, nact, vp(dCmm0.ptr), n, vp(dCpp.ptr), m,
                                     vp(dCpm.ptr), m, vp(dR.ptr), m, vp(dDx.ptr), nact,
                                     vp(dCee.ptr), m, vp(dCvv.ptr), nact, vp(dTmp.ptr), nact,
                                     None), "moao_hip_cee_cvv_dev")
    synchronize()
    out = dCee.to_host()
    for d in (dCpp, dCpm, dR, dDx, dCee, dCvv, dTmp):
        d.free()
    return out


def cee_cvv_flops(n, m, nact):
    """syr2k + symm + gemm + symm + gemm of lapack/compute/cee_cvv.cpp:55-117."""
    return 2.0 * m * m * n + 2.0 * m * n * n + 2.0 * m * m * n + 2.0 * nact * m * m \
        + 2.0 * nact * nact * m


def stages(tomo, n, m, dCmm0, dCtm0, dR, ndir=8, N=2048, dm_nssp=79, reps=2):
    """Observation-loop stages on the same system: Cee/Cvv (nact ~ 5000 on an 80 x 80 DM grid)
    and the intersample PSF (N x N FP32 FFTs, ndir directions batched).  Inputs resident in
    HBM; each timed over `reps` calls after one warm-up."""
    import ctypes
    from moao_amd import psf_inputs, systems
    from moao_amd.device import DeviceArray, Profile, synchronize
    from moao_amd.lib import check, lib
    from moao_amd.matcov import CTT, matcov_tile_dev
    vp = ctypes.c_void_p
    dm = psf_inputs.make_dm(dm_nssp)
    nact = dm.nact
    out = {}
    # ---- Cee/Cvv
    dDx = DeviceArray.from_host(systems.dx_matrix(nact, m, n))
    dCtt = DeviceArray((m, m))
    matcov_tile_dev(dCtt.ptr, m, m, 0, 0, m, tomo, CTT)
    dCpp, dCpm = DeviceArray((m, m)), DeviceArray((m, n))
    dCee, dCvv, dTmp = DeviceArray((m, m)), DeviceArray((nact, nact)), DeviceArray((nact, m))
    ms = []
    for it in range(reps + 1):
        dCpp.copy_from(dCtt)   # Cpp and Cpm are clobbered by each call, as in the reference
        dCpm.copy_from(dCtm0)
        synchronize()
        with Profile() as prof:
            t0 = time.perf_counter()
            check(lib().moao_hip_cee_cvv_dev(n, m, nact, vp(dCmm0.ptr), n, vp(dCpp.ptr), m,
                                             vp(dCpm.ptr), m, vp(dR.ptr), m, vp(dDx.ptr), nact,
                                             vp(dCee.ptr), m, vp(dCvv.ptr), nact, vp(dTmp.ptr),
                                             nact, None), "moao_hip_cee_cvv_dev")
            synchronize()
            dt = time.perf_counter() - t0
        if it:
            ms.append(dt * 1e3)
    g = prof.result["gemm_f64"]
    fl = cee_cvv_flops(n, m, nact)
    out["cee_cvv"] = {"n": n, "m": m, "nact": nact, "ms": min(ms), "flops": fl,
                      "tflops": fl / (min(ms) * 1e-3) / 1e12,
                      "gemm_tflops": g["work"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] else None,
                      "peak": PEAK_FP64 / 1e12, "bound": "mfma (FP64)"}
    for d in (dDx, dCtt, dCpp, dCpm, dCee, dTmp):
        d.free()
    # ---- intersample PSF on the Cvv just computed, ndir directions batched
    idx = psf_inputs.make_idx(dm)
    diam = tomo.view().diam
    a2f = psf_inputs.make_abs2fi(N, diam, dm)
    otf = psf_inputs.make_otf_tel(N, diam, 0.28)
    h = ctypes.c_void_p()
    lp = ctypes.POINTER(ctypes.c_long)
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().moao_hip_intersample_create(idx.size, dm.ndiam, N, idx.ctypes.data_as(lp),
                                            a2f.ctypes.data_as(dp), otf.ctypes.data_as(dp), 0,
                                            ctypes.byref(h)), "intersample_create")
    del idx
    dC = DeviceArray((ndir * nact * nact,))
    for k in range(ndir):
        lib().moao_hip_memcpy_d2d(vp(dC.ptr + k * nact * nact * 8), vp(dCvv.ptr),
                                  nact * nact * 8)
    dP = DeviceArray((ndir * N * N,))
    ms = []
    for it in range(reps + 1):
        synchronize()
        with Profile() as prof:
            t0 = time.perf_counter()
            check(lib().moao_hip_intersample_process_dev(h, vp(dC.ptr), ndir, vp(dP.ptr), None),
                  "intersample_process_dev")
            synchronize()
            dt = time.perf_counter() - t0
        if it:
            ms.append(dt * 1e3)
    f = prof.result["fft"]
    # algorithmic HBM bytes of one direction: Cvv gather (8 B per entry + 4 B index) plus
    # half of 3 complex 2-D FFTs of N x N FP32 (two directions share each transform; each
    # 2 row passes + 1 transpose, read + write 8 B), half of the pointwise passes
    # (~6 N^2 x 8 B per pair) and the FP64 PSF write
    byt = ndir * (12.0 * nact * nact + 1.5 * 3 * 2 * 8.0 * N * N + 3 * 8.0 * N * N
                  + 8.0 * N * N)
    psf = dP.to_host()[:N * N].reshape(N, N)
    out["psf"] = {"N": N, "directions": ndir, "nact": nact, "precision": "fp32 FFT",
                  "ms": min(ms), "ms_per_direction": min(ms) / ndir,
                  "fft_kernel_ms": f["ms"], "achieved_GBps": byt / (min(ms) * 1e-3) / 1e9,
                  "peak_GBps": 8000.0, "bound": "hbm",
                  "strehl_dir0": float(psf[N // 2, N // 2])}
    lib().moao_hip_intersample_destroy(h)
    dC.free()
    dP.free()
    dCvv.free()
    return out


def obs_prepare(d, maxrefine, maxobs, ndir, N, dm_nssp):
    """Files of the observation-loop system in directory d (system, maxrefine * maxobs + 1
    atmosphere profiles, Dx, idx, abs2fi, otftel) and the `moao` command line that runs it."""
    import dataclasses
    from moao_amd import fits, psf_inputs, systems
    from moao_amd.tomo import Tomo
    spec = dataclasses.replace(systems.WORKLOADS["elt"], ntarget=ndir)
    s, atms = systems.make_system(d, spec, nprof=maxrefine * maxobs + 1)
    t = Tomo(s, atms[0])
    n, m = t.nmeas(), t.nmeasts()
    dm = psf_inputs.make_dm(dm_nssp)
    fits.write_image(os.path.join(d, "Dx.fits"), systems.dx_matrix(dm.nact, m, n), dm.nact, m)
    idx = psf_inputs.make_idx(dm)
    a2f = psf_inputs.make_abs2fi(N, spec.diam, dm)
    otf = psf_inputs.make_otf_tel(N, spec.diam, spec.obs)
    fits.write_image(os.path.join(d, "idx.fits"), idx, idx.size, 1)
    fits.write_image(os.path.join(d, "abs2fi.fits"), a2f, N, N)
    fits.write_image(os.path.join(d, "otftel.fits"), otf, N, N)
    exe = os.path.join(ROOT, "moao_amd", "_lib", "moao")
    cmd = [exe, "--sys_path", d, "--atm_path", d, "--maxrefine", str(maxrefine), "--maxobs",
           str(maxobs), "--snapPerNight", str(maxobs), "--psf", "1", "--out", d,
           "--write_short", "0", "--dm_np", str(dm.ndiam)]
    return n, m, dm, idx, a2f, otf, cmd


def observation_loop(maxrefine=4, maxobs=16, ndir=8, N=2048, dm_nssp=79):
    """The reference's refine / observation loop (lapack/moao_lapack.cpp:95-188) through the
    `moao` driver on the n = 18,688 system (40x40 subapertures, 6 LGS + 2 NGS, 10 layers) with
    8 truth directions: maxrefine R-solves (all directions at once) and maxobs observation steps
    per refine (profiles prof{r*maxobs-1} and prof{r*maxobs+o}), each step Cmm, then per
    direction Ctm, Ctt and Cee/Cvv (80x80 DM, Gaussian Dx), the 8 PSFs (N x N, FP32 FFTs) in
    one batched call and the long-exposure PSF of each direction.  Wall time of the whole run,
    the driver's own stage times, and the Strehl of the last step's direction-0 PSF with FP32
    and FP64 transforms."""
    import ctypes
    import subprocess
    from moao_amd import fits
    from moao_amd.lib import check, lib
    d = tempfile.mkdtemp(prefix="moao_obs_")
    n, m, dm, idx, a2f, otf, cmd = obs_prepare(d, maxrefine, maxobs, ndir, N, dm_nssp)
    nact = dm.nact
    t0 = time.perf_counter()
    r = subprocess.run(cmd, capture_output=True, text=True)
    wall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError("moao observation loop failed: " + r.stdout[-2000:] + r.stderr[-2000:])
    times = {}
    for ln in r.stdout.splitlines():
        parts = ln.split(":")
        if len(parts) == 2 and parts[0].strip() in ("preprocess", "matcov", "R", "Cee_Cvv",
                                                    "psf", "total"):
            times[parts[0].strip()] = float(parts[1])
    Cvv, _ = fits.read_image(os.path.join(d, "Cvv.fits"))
    strehl = {}
    for prec in (0, 1):
        h = ctypes.c_void_p()
        lp_ = ctypes.POINTER(ctypes.c_long)
        dp = ctypes.POINTER(ctypes.c_double)
        check(lib().moao_hip_intersample_create(idx.size, dm.ndiam, N, idx.ctypes.data_as(lp_),
                                                a2f.ctypes.data_as(dp), otf.ctypes.data_as(dp),
                                                prec, ctypes.byref(h)), "intersample_create")
        psf = np.empty(N * N)
        check(lib().moao_hip_intersample_process(h, Cvv.ctypes.data_as(dp),
                                                 psf.ctypes.data_as(dp)), "intersample_process")
        lib().moao_hip_intersample_destroy(h)
        strehl["fp32_fft" if prec == 0 else "fp64_fft"] = float(psf.reshape(N, N)[N // 2, N // 2])
    longexp, _ = fits.read_image(os.path.join(d, f"psf_nbrefine{maxrefine}_maxobs{maxobs}_gal0.fits"))
    return {"n": n, "m": m, "nact": nact, "directions": ndir, "N": N, "maxrefine": maxrefine,
            "maxobs": maxobs, "wall_s": wall, "driver_times_s": times,
            "strehl_last_step_dir0": strehl,
            "strehl_long_exposure_dir0": float(longexp.reshape(N, N)[N // 2, N // 2]),
            "_inputs": (d, idx, a2f, otf, dm.ndiam)}


def serving(nreq, mode, slots, seed=2024, util=0.8, solve_factor=1.0):
    """Request stream of uneven systems with Poisson arrivals through the serving engine.

    Each request draws n_wfs ~ U{4..9} measuring sensors (two thirds LGS, plus the truth
    sensor) and Nssp ~ U{20..74} on the D=39 m telescope (n = 2 n_wfs Nvalid, redrawn while
    n > 40,000), 10 layers with fresh
    seeded Cn2 and r0 ~ U(0.08, 0.2) m.  Arrivals are Poisson at a rate that would keep one
    GPU ~util busy by a service-time model (solve flops, matcov elements, R bytes).  A request
    is one refine step: Tomo from files, matcov Cmm and Ctm, solve, R back to the host."""
    import threading
    from moao_amd import systems
    from moao_amd.lib import MOAO_FP64, MOAO_MIXED_BF16, MOAO_MIXED_FP32
    from moao_amd.serve import Server
    rng = np.random.default_rng(seed + 1)
    tmp = tempfile.mkdtemp(prefix="moao_serve_")
    reqs = systems.request_stream(tmp, nreq, seed)
    # service-time model of one request on an otherwise idle GPU: an FP64 solve at 40 TFLOP/s
    # (a mixed one `solve_factor` times that: FP64 residuals on top of the low-precision
    # solves), matcov at 2.3e9 elements/s (10 layers), R to pageable host memory at 20 GB/s
    est = [solve_factor * solve_flops(n, m) / 40e12 + (n * (n + 1) / 2 + m * n) / 2.3e9
           + 8.0 * m * n / 20e9 for _, _, n, m in reqs]
    rate = util / (sum(est) / len(est))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    md = {"fp64": MOAO_FP64, "mixed": MOAO_MIXED_FP32, "bf16": MOAO_MIXED_BF16}[mode]
    results = [None] * nreq
    with Server(slots=slots) as srv:
        tickets = [None] * nreq
        t0 = time.perf_counter()

        done = [0]
        lock = threading.Lock()

        def waiter(k):
            results[k] = srv.wait(tickets[k], raise_on_error=False)
            with lock:
                done[0] += 1
                if done[0] % 32 == 0:
                    print(f"[bench] serving {mode}: {done[0]}/{nreq} done "
                          f"({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)

        threads = []
        for k, (s, a, n, m) in enumerate(reqs):
            dt = arrivals[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets[k] = srv.submit(s, a, 0, md, tol=1e-9, max_iter=6, want_R=True,
                                    n_m=(n, m))
            th = threading.Thread(target=waiter, args=(k,))
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
    ok = [r for r in results if r["status"] == 0]
    lat = sorted(r["latency_s"] * 1e3 for r in ok)
    q = sorted(r["queue_s"] * 1e3 for r in ok)
    ns = [n for _, _, n, _ in reqs]

    def pct(v, p):
        return v[min(len(v) - 1, int(math.ceil(p * len(v))) - 1)] if v else None

    return {"requests": nreq, "completed": len(ok), "mode": mode, "slots": slots,
            "arrival_rate_per_s": rate, "wall_s": wall,
            "throughput_solves_per_s": len(ok) / wall,
            "latency_ms": {"p50": pct(lat, 0.5), "p99": pct(lat, 0.99),
                           "max": lat[-1] if lat else None},
            "queue_ms": {"p50": pct(q, 0.5), "p99": pct(q, 0.99)},
            "peak_running": st["peak_running"], "peak_bytes": st["peak_bytes"],
            "graph_captures": st["graph_captures"], "graph_replays": st["graph_replays"],
            "n": {"min": min(ns), "mean": sum(ns) / len(ns), "max": max(ns)},
            "note": "latency = submit -> R on the host (queueing + Tomo + matcov + solve)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="elt74")
    ap.add_argument("--modes", default="fp64,mixed,bf16",
                    help="fp64, mixed (FP32 MFMA updates), bf16 (split-BF16 updates)")
    ap.add_argument("--tol", type=float, default=3e-10,
                    help="mixed path: relative residual ||Ctm - R Cmm||/||Ctm|| to stop at "
                         "(3e-10 keeps the relative error of R below 1e-6 at n=40k, where "
                         "error / residual ~ 1.4e3)")
    ap.add_argument("--max-iter", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-matcov", action="store_true")
    ap.add_argument("--obs-loop", type=int, default=1,
                    help="run the refine/observation loop leg (n=18688, 8 directions, PSFs)")
    ap.add_argument("--serve", type=int, default=256,
                    help="requests in the serving leg (0 = skip it)")
    ap.add_argument("--serve-modes", default="fp64",
                    help="serving legs to run (fp64, mixed, bf16); the split-BF16 leg after the "
                         "FP64 one stopped completing requests once and is not run by default")
    ap.add_argument("--serve-mixed", type=int, default=128,
                    help="requests in the mixed-precision serving legs (their solves take "
                         "about three times longer; 0 = as many as --serve)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--no-stages", action="store_true",
                    help="skip the Cee/Cvv and PSF stage measurements")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the bench, run 3 plain device-to-device copies of this many "
                         "MiB (the calibration kernel for HBM byte counters)")
    ap.add_argument("--obs-prepare", metavar="DIR",
                    help="only write the observation-loop system (--obs-refine x --obs-steps "
                         "steps) to DIR and print the moao command that runs it (no GPU use)")
    ap.add_argument("--obs-refine", type=int, default=4)
    ap.add_argument("--obs-steps", type=int, default=16)
    args = ap.parse_args()
    if args.obs_prepare:
        os.makedirs(args.obs_prepare, exist_ok=True)
        *_, cmd = obs_prepare(args.obs_prepare, args.obs_refine, args.obs_steps, 8, 2048, 79)
        print(" ".join(cmd))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)

    from moao_amd import systems
    from moao_amd.device import DeviceArray, Profile, synchronize
    from moao_amd.lib import MOAO_MIXED_BF16, MOAO_MIXED_FP32, check, lib
    from moao_amd.matcov import CMM, CTM, matcov_tile_dev
    from moao_amd.reconstructor import reconstructor_dev, reconstructor_mixed_dev
    from moao_amd.tomo import Tomo

    check(lib().moao_hip_init(0), "moao_hip_init")
    if args.copy_probe_mib > 0:
        nd = args.copy_probe_mib * (1 << 20) // 8
        a, b = DeviceArray((nd,)), DeviceArray((nd,))
        for _ in range(3):
            b.copy_from(a)
        synchronize()
        print(f"[bench] copy probe: 3 x {args.copy_probe_mib} MiB device-to-device",
              file=sys.stderr, flush=True)
        a.free()
        b.free()
    tmp = tempfile.mkdtemp(prefix="moao_bench_")
    sys_file, atms = systems.make_system(tmp, args.workload)
    tomo = Tomo(sys_file, atms[0])
    n, m = tomo.nmeas(), tomo.nmeasts()
    print(f"[bench] workload {args.workload}: n={n} m={m}", file=sys.stderr, flush=True)

    dCmm0 = DeviceArray((n, n))
    dCtm0 = DeviceArray((m, n))
    t0 = time.perf_counter()
    with Profile() as prof_mc:
        matcov_tile_dev(dCmm0.ptr, n, n, 0, 0, n, tomo, CMM)
        tomo.set_target(0)
        matcov_tile_dev(dCtm0.ptr, m, n, 0, 0, m, tomo, CTM)
        synchronize()
    t_matcov = time.perf_counter() - t0
    mc = prof_mc.result["matcov"]
    print(f"[bench] matcov {t_matcov:.2f} s ({mc['ms']:.1f} ms kernel)", file=sys.stderr,
          flush=True)

    dCmm, dCtm = DeviceArray((n, n)), DeviceArray((m, n))
    results = {}
    flops = solve_flops(n, m)
    Rs = {}
    for mode in args.modes.split(","):
        extra = {}

        def step():
            if mode == "fp64":
                dCmm.copy_from(dCmm0)
                dCtm.copy_from(dCtm0)
                info = reconstructor_dev(n, m, dCmm.ptr, n, dCtm.ptr, m, 1)
                if info != 0:
                    raise RuntimeError(f"reconstructor info={info}")
            else:  # Cmm is read only on the mixed paths
                dCtm.copy_from(dCtm0)
                md = MOAO_MIXED_FP32 if mode == "mixed" else MOAO_MIXED_BF16
                rc, res, its = reconstructor_mixed_dev(n, m, dCmm0.ptr, n, dCtm.ptr, m, tomo,
                                                       args.tol, args.max_iter,
                                                       allow_noconv=True, mode=md)
                extra.update(status=rc, resid=res, iters=its)
            synchronize()
        for _ in range(args.warmup):
            step()
        synchronize()
        lat = []
        with Profile() as prof:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                s0 = time.perf_counter()
                step()
                lat.append(time.perf_counter() - s0)
            synchronize()
            total = time.perf_counter() - t0
        results[mode] = {"total": total, "lat": lat, "prof": prof.result, **extra}
        Rs[mode] = dCtm.to_host()
        if mode == "fp64":  # the FP64 R feeds the observation-loop stages below
            dR64 = DeviceArray((m, n))
            dR64.copy_from(dCtm)
        print(f"[bench] {mode}: {total / args.steps * 1e3:.1f} ms/solve {extra}",
              file=sys.stderr, flush=True)

    for k in Rs:
        if k != "fp64" and "fp64" in Rs:
            e = float(np.linalg.norm(Rs[k] - Rs["fp64"]) / np.linalg.norm(Rs["fp64"]))
            results[k]["rel_err_vs_fp64"] = e
            print(f"[bench] {k} vs fp64 relative Frobenius error {e:.3e}", file=sys.stderr,
                  flush=True)
    if "fp64" in Rs and len(Rs) > 1:
        # Cee = Ctt - Ctm R^T - R Ctm^T + R Cmm R^T of each mixed R against the FP64 R's
        cee64 = cee_of(tomo, n, m, dCmm0, dCtm0, Rs["fp64"])
        for k in Rs:
            if k != "fp64":
                ck = cee_of(tomo, n, m, dCmm0, dCtm0, Rs[k])
                results[k]["cee_rel_err_vs_fp64"] = float(np.linalg.norm(ck - cee64) /
                                                          np.linalg.norm(cee64))
                print(f"[bench] {k}: Cee relative error vs fp64 "
                      f"{results[k]['cee_rel_err_vs_fp64']:.3e}", file=sys.stderr, flush=True)
    del Rs
    # headline: the fastest path whose R meets the 1e-6 bound (FP64 is the reference path)
    ok = [k for k in results
          if k == "fp64" or (results[k].get("rel_err_vs_fp64") is not None
                             and results[k]["rel_err_vs_fp64"] <= 1e-6
                             and results[k].get("status") == 0)]
    # (a run of one mixed mode alone has no FP64 R to check against: it heads its own line)
    head = min(ok, key=lambda k: results[k]["total"]) if ok else next(iter(results))
    r = results[head]
    ms_step = r["total"] / args.steps * 1e3
    g = results["fp64"]["prof"]["gemm_f64"] if "fp64" in results else r["prof"]["gemm_f64"]
    achieved = g["work"] / (g["ms"] * 1e-3) if g["ms"] > 0 else 0.0
    lat = sorted(r["lat"])
    out = {
        "metric": METRIC,
        "value": args.steps / r["total"],
        "unit": "R-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32+f64", "bf16": "bf16+f64"}[head],
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: seeded D=39 m system, n={n} meas, m={m} truth "
                               f"meas, {tomo.dims()['nLayer']} layers, one R-solve per step",
                   "n": n, "m": m, "path": head},
        "latency_ms": {"p50": 1e3 * statistics.median(lat),
                       "p99": 1e3 * lat[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)]},
        "solve_tflops": flops / (ms_step * 1e-3) / 1e12,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                     "traffic": hbm_traffic(args.workload),
                     "kernel": "FP64 GEMM engine: gemm_dl_kernel + gemm_kernel (all launches of the timed solves)",
                     "launches": g["count"], "avg_launch_ms": g["ms"] / max(1, g["count"])},
        "stages_ms_per_solve": {k: v["ms"] / args.steps for k, v in r["prof"].items()
                                if v["count"]},
        "matcov_s": t_matcov,
    }
    for mode, rr in results.items():
        out[f"{mode}_solves_per_s"] = args.steps / rr["total"]
        lt = sorted(rr["lat"])
        out[f"{mode}_latency_ms"] = {"p50": 1e3 * statistics.median(lt), "p99": 1e3 * lt[-1]}
    # the low-precision GEMM class: FP32 MFMA (mixed) or the split-BF16 kernel (bf16, three
    # BF16 MFMAs per FP32-equivalent product: its ceiling is a third of the BF16 peak)
    lowpeak = {"mixed": PEAK_FP32, "bf16": PEAK_BF16 / 3.0}
    for mk in ("mixed", "bf16"):
        if mk not in results:
            continue
        mr = results[mk]
        out.setdefault("mixed", {})["fp32" if mk == "mixed" else "bf16"] = {
            "rel_err_vs_fp64": mr.get("rel_err_vs_fp64"),
            "cee_rel_err_vs_fp64": mr.get("cee_rel_err_vs_fp64"), "resid": mr.get("resid"),
            "refinement_steps": mr.get("iters"), "status": mr.get("status"),
            "speedup_vs_fp64": (results["fp64"]["total"] / mr["total"]
                                if "fp64" in results else None),
            "stages_ms_per_solve": {k: v["ms"] / args.steps
                                    for k, v in mr["prof"].items() if v["count"]},
            # MFMA rate of each GEMM class: algorithmic (FP32-equivalent) flops / summed time
            "gemm_tflops": {k: {"achieved": v["work"] / (v["ms"] * 1e-3) / 1e12,
                                "peak": (PEAK_FP64 if k == "gemm_f64" else lowpeak[mk]) / 1e12}
                            for k, v in mr["prof"].items()
                            if k.startswith("gemm_") and v["count"] and v["ms"] > 0}}
    if not args.no_stages and "fp64" in results:
        # R of the last FP64 step is in dR64; Cmm0 is the regenerated Cmm
        out["stages"] = stages(tomo, n, m, dCmm0, dCtm0, dR64)
        L = tomo.dims()["nLayer"]
        elems = n * (n + 1) / 2 + m * n
        mfl = (30 + 115 * L) * n * (n + 1) / 2 + (30 + 103 * L) * m * n  # chameleon/flops.h
        out["stages"]["matcov"] = {
            "elements": elems, "ms": mc["ms"], "elements_per_s": elems / (mc["ms"] * 1e-3),
            "GBps_written": 8 * elems / (mc["ms"] * 1e-3) / 1e9,
            "model_flops": mfl, "model_tflops": mfl / (mc["ms"] * 1e-3) / 1e12,
            "bound": "fp64 valu (transcendentals)", "peak_fp64_vector_tflops": 78.6}
        print(f"[bench] stages: {json.dumps(out['stages'])}", file=sys.stderr, flush=True)
    if args.serve > 0:
        out["serving"] = []
        for sm in args.serve_modes.split(","):
            # service-time factor of a mixed request against FP64, from the timed solves above
            fac = 1.0
            if sm != "fp64" and sm in results and "fp64" in results:
                fac = results[sm]["total"] / results["fp64"]["total"]
            elif sm != "fp64":
                fac = 4.0
            nreq = args.serve if sm == "fp64" or args.serve_mixed <= 0 else args.serve_mixed
            sv = serving(nreq, sm, args.slots, solve_factor=fac)
            print(f"[bench] serving {sm}: {sv['throughput_solves_per_s']:.2f} solves/s, "
                  f"p50 {sv['latency_ms']['p50']:.0f} ms, p99 {sv['latency_ms']['p99']:.0f} ms",
                  file=sys.stderr, flush=True)
            out["serving"].append(sv)
    obs = None
    if args.obs_loop:
        print("[bench] observation loop (n=18688, 4 refine x 16 obs, 8 directions)",
              file=sys.stderr, flush=True)
        obs = observation_loop()
        out["observation_loop"] = obs
        print(f"[bench] observation loop: {obs['wall_s']:.1f} s, Strehl "
              f"{obs['strehl_last_step_dir0']}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        print("[bench] cpu baseline: full oracle solve", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(dCmm0, dCtm0, n, m)
        if not args.no_cpu_matcov:
            print("[bench] cpu baseline: oracle matcov sample", file=sys.stderr, flush=True)
            cm = cpu_matcov(sys_file, atms[0], n, m)
            out["cpu_baseline"]["matcov"] = cm
            # end to end (matcov + solve) on the host against the GPU's matcov + FP64 solve
            solve_s = 1.0 / out["cpu_baseline"]["value"]
            out["cpu_baseline"]["end_to_end_s"] = {
                "threaded_matcov": cm["threaded"]["full_s"] + solve_s,
                "serial_matcov": cm["serial"]["full_s"] + solve_s}
            if "fp64" in results:
                out["end_to_end_s"] = {"gpu": t_matcov + results["fp64"]["total"] / args.steps}
        if obs is not None:
            # the last observation step's direction-0 PSF on the host: the oracle's FP64
            # restatement of intersample_process (numpy FFTs), timed, and the Strehl difference
            from moao_amd import fits
            from oracle import intersample as oi
            d, idx, a2f, otf, ndiam = obs.pop("_inputs")
            Cvv, _ = fits.read_image(os.path.join(d, "Cvv.fits"))
            t0 = time.perf_counter()
            P = oi.process(Cvv, idx, a2f, otf, obs["N"], ndiam)
            dt = time.perf_counter() - t0
            st = float(P[obs["N"] // 2, obs["N"] // 2])
            out["cpu_baseline"]["psf"] = {
                "s_per_direction": dt, "strehl_dir0": st,
                "sample": f"oracle numpy FP64 intersample of the last step's direction-0 Cvv, "
                          f"N={obs['N']}"}
            obs["strehl_delta_vs_cpu_fp64"] = {
                k: v - st for k, v in obs["strehl_last_step_dir0"].items()}
    if obs is not None:
        obs.pop("_inputs", None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
