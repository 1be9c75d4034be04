"""The `moao` command-line driver (the reference's refine/observation loop over the C-ABI,
lapack/moao_lapack.cpp:30-188): usage on the CPU; on the GPU one refine + one observation step
with Dx.fits and the PSF inputs, checked against the C-ABI path (R, bit for bit) and the
oracle (Cee/Cvv closed form <= 1e-10, PSF <= 2e-5 of the peak with FP32 FFTs)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "moao_amd", "_lib", "moao")


def test_usage_without_paths():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--sys_path" in r.stderr


@pytest.mark.gpu
def test_refine_and_observation_step(tmp_path):
    from moao_amd import fits, psf_inputs, systems
    from moao_amd import matcov as gm
    from moao_amd.reconstructor import reconstructor
    from moao_amd.tomo import Tomo
    from oracle import intersample as oi
    from oracle import lapack_path as lp
    from oracle.check import rel_fro
    d = str(tmp_path)
    s, atms = systems.make_system(d, "small", nprof=2)    # prof-1 (refine), prof0 (observe)
    t = Tomo(s, atms[0])
    n, m = t.nmeas(), t.nmeasts()
    dm = psf_inputs.make_dm(10)
    nact = dm.nact
    Dx = systems.dx_matrix(nact, m, n)
    fits.write_image(os.path.join(d, "Dx.fits"), Dx, nact, m)
    idx = psf_inputs.make_idx(dm)
    N = oi.default_n(dm.ndiam)
    a2f = psf_inputs.make_abs2fi(N, 8.0, dm)
    otf = psf_inputs.make_otf_tel(N, 8.0, 0.28)
    fits.write_image(os.path.join(d, "idx.fits"), idx, idx.size, 1)
    fits.write_image(os.path.join(d, "abs2fi.fits"), a2f, N, N)
    fits.write_image(os.path.join(d, "otftel.fits"), otf, N, N)
    print(f"[cli] n={n} m={m} nact={nact} N={N}", flush=True)
    r = subprocess.run([EXE, "--sys_path", d, "--atm_path", d, "--maxrefine", "1", "--maxobs",
                        "1", "--psf", "1", "--out", d], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # refine step (prof-1): R through the C-ABI path, bit for bit
    t.set_target(0)
    C, B = gm.cmm(t), gm.ctm(t)
    Rg = B.copy(order="F")
    assert reconstructor(n, m, C.copy(order="F"), n, Rg, m, 1) == 0
    Rf, dims = fits.read_image(os.path.join(d, "R.fits"))
    assert dims == (m, n) and np.array_equal(Rf.reshape((m, n), order="F"), Rg)
    # observation step (prof0): Cee/Cvv against the oracle chain on the same matrices
    t.update_atm(atms[1])
    t.set_target(0)
    C, B, T = gm.cmm(t), gm.ctm(t), gm.ctt(t)
    Cee, Cvv = lp.cee_cvv(C.copy(order="F"), T.copy(order="F"), B.copy(order="F"), Rg,
                          np.asfortranarray(Dx))
    Ceef, _ = fits.read_image(os.path.join(d, "Cee.fits"))
    Cvvf, _ = fits.read_image(os.path.join(d, "Cvv.fits"))
    assert rel_fro(Ceef.reshape((m, m), order="F"), Cee) <= 1e-10
    assert rel_fro(Cvvf.reshape((nact, nact), order="F"), Cvv) <= 1e-10
    psf, pd = fits.read_image(os.path.join(d, "psf_oloop0_iloop0_gal0.fits"))
    ref = oi.process(Cvvf, idx, a2f, otf, N, dm.ndiam)
    assert pd == (N, N)
    assert np.max(np.abs(psf.reshape(N, N) - ref)) <= 2e-5 * np.max(np.abs(ref))
