"""The header-only MOAO_HIP:: layer compiles against a Tomo_struct-shaped caller and links
against libmoao_hip.so with plain g++ (CPU: link check only; GPU: the refine step through it
gives the same R, bit for bit, as the C-ABI path and matches the oracle's LAPACK path; the
reference's intersample sequence through it gives the oracle's PSF).  Where the reference tree
is present (the build container), a caller is also compiled against the reference's own
Tomo_struct<double> header and linked with the hand-built reference library."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

LIBDIR = os.path.join(ROOT, "moao_amd", "_lib")


def _build(tmp_path):
    exe = str(tmp_path / "dropin")
    cmd = ["g++", "-std=c++11", "-O2", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "cpp", "dropin.cpp"), "-o", exe, f"-L{LIBDIR}",
           "-lmoao_hip", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_dropin_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "--link-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "gfx950" in r.stdout


@pytest.mark.gpu
def test_dropin_refine_step_matches_c_abi_and_oracle(tmp_path):
    from moao_amd import matcov as gm
    from moao_amd import systems
    from moao_amd.reconstructor import reconstructor
    from moao_amd.tomo import Tomo
    from oracle import lapack_path as lp
    from oracle.check import rel_fro
    exe = _build(tmp_path)
    s, atms = systems.make_system(str(tmp_path), "small")
    out = str(tmp_path / "R.bin")
    print("[dropin] small workload through MOAO_HIP::", flush=True)
    r = subprocess.run([exe, s, atms[0], out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    t = Tomo(s, atms[0])
    t.set_target(0)
    C = gm.cmm(t)
    B = gm.ctm(t)
    n, m = C.shape[0], B.shape[0]
    Rd = np.fromfile(out).reshape((m, n), order="F")
    Cg, Rg = C.copy(order="F"), B.copy(order="F")
    assert reconstructor(n, m, Cg, n, Rg, m, 1) == 0
    assert np.array_equal(Rd, Rg)
    Co, Ro = C.copy(order="F"), B.copy(order="F")
    assert lp.reconstructor(Co, Ro) == 0
    assert rel_fro(Rd, Ro) <= 1e-10


REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "common", "tomo_struct")),
                    reason="the reference tree is only present in the build container")
def test_dropin_compiles_against_the_reference_tomo_struct(tmp_path):
    """MOAO_HIP:: instantiated with the reference's real Tomo_struct<double> (its header, its
    compiled implementation from oracle/_ref) for every call the reference's lapack driver
    makes, FP64 and FP32 reconstructor included.  Compile and link only: nothing runs."""
    refso = os.path.join(ROOT, "oracle", "_ref", "libmoao_ref.so")
    if not os.path.exists(refso):
        subprocess.run(["make", "-f", os.path.join("oracle", "Makefile.ref")], cwd=ROOT,
                       check=True, capture_output=True)
    exe = str(tmp_path / "dropin_ref")
    cmd = ["g++", "-std=c++11", "-O1", "-w", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(REF, 'common')}",
           os.path.join(ROOT, "tests", "cpp", "dropin_ref_tomo.cpp"), "-o", exe,
           f"-L{LIBDIR}", "-lmoao_hip", refso, f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.gpu
@pytest.mark.parametrize("cvv_type", ["f64", "f32"])
def test_dropin_intersample_sequence_matches_oracle(tmp_path, cvv_type):
    """prepare/init/process/free with the reference's signatures on FITS inputs in the
    reference's layout; the PSF left in isample.dphi equals the oracle's within the FP64-FFT
    bar (1e-11 of the peak; a float Cvv rounds the input first, compared with the oracle on
    the same rounded values)."""
    from moao_amd import fits
    from moao_amd import psf_inputs as pi
    from oracle import intersample as oi
    exe = _build(tmp_path)
    dm = pi.make_dm(10, 0.2, 0.28)
    np_ = dm.ndiam
    N = oi.default_n(np_)
    idx = pi.make_idx(dm)
    a2f = pi.make_abs2fi(N, 8.0, dm)
    otf = pi.make_otf_tel(N, 8.0, 0.28)
    fits.write_image(str(tmp_path / "idx.fits"), idx, idx.size, 1)
    fits.write_image(str(tmp_path / "abs2fi.fits"), a2f, N, N)
    fits.write_image(str(tmp_path / "otftel.fits"), otf, N, N)
    rng = np.random.default_rng(12)
    G = rng.standard_normal((dm.nact, dm.nact))
    cvv = (G @ G.T / dm.nact * 3e-4).flatten(order="F")
    cvv.tofile(str(tmp_path / "Cvv.bin"))
    out = str(tmp_path / "psf.bin")
    args = [exe, "--psf", str(tmp_path), str(dm.nact), str(np_), str(tmp_path / "Cvv.bin"), out]
    if cvv_type == "f32":
        args.append("f32")
        cvv = cvv.astype(np.float32).astype(np.float64)
    print(f"[dropin] intersample sequence, Cvv {cvv_type}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"N={N}" in r.stdout
    psf = np.fromfile(out).reshape(N, N)
    ref = oi.process(cvv, idx, a2f, otf, N, np_)
    err = np.max(np.abs(psf - ref)) / np.max(np.abs(ref))
    print(f"    err={err:.2e} strehl={ref[N // 2, N // 2]:.6f}", flush=True)
    assert err <= 1e-11
