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


def _psf_system(d, nprof=2):
    from moao_amd import fits, psf_inputs, systems
    from moao_amd.tomo import Tomo
    from oracle import intersample as oi
    s, atms = systems.make_system(d, "small", nprof=nprof)
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
    return s, atms, t, n, m, dm, Dx, idx, a2f, otf, N


def test_check_usage():
    r = subprocess.run([os.path.join(ROOT, "moao_amd", "_lib", "check_moao")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--sys_check" in r.stderr


@pytest.mark.gpu
def test_check_mode_table_and_gate(tmp_path):
    """check_moao (lapack/check_moao_lapack.cpp): every stage of one pass compared with
    reference files -- here written by the oracle chain (matcov restatement, scipy LAPACK path,
    Cee/Cvv closed form, numpy PSF) -- by compareMatrices.  The table's maximum errors and
    Frobenius norms equal the oracle's own compare_matrices on the same matrices (the check
    harness measures what it says), the gate passes at an accuracy the stages meet and fails,
    with exit status 1, once one reference file is wrong."""
    import re
    from moao_amd import fits
    from oracle import intersample as oi
    from oracle import lapack_path as lp
    from oracle import matcov as om
    from oracle.check import compare_matrices
    from oracle.tomo import Tomo as OTomo
    d = str(tmp_path)
    s, atms, t, n, m, dm, Dx, idx, a2f, otf, N = _psf_system(d)
    chk = os.path.join(d, "check")
    os.makedirs(chk)
    o = OTomo.from_files(s, atms[0])
    o.set_target(0)
    Co, Bo, To = om.cmm(o), om.ctm(o), om.ctt(o)
    Ro = Bo.copy(order="F")
    assert lp.reconstructor(Co.copy(order="F"), Ro) == 0
    Ceo, Cvo = lp.cee_cvv(Co.copy(order="F"), To.copy(order="F"), Bo.copy(order="F"), Ro,
                          np.asfortranarray(Dx))
    Po = oi.process(Cvo.flatten(order="F"), idx, a2f, otf, N, dm.ndiam)
    # each reference flat in the library's memory order: column-major matrices; the PSF in
    # the layout of the reference's dphi (what oracle.intersample.process returns C-ordered)
    refs = {k: (np.asarray(v).flatten(order="F"), v.shape[0], v.shape[1])
            for k, v in {"Cmm": Co, "Ctm": Bo, "R": Ro, "Ctt": To, "Cee": Ceo,
                         "Cvv": Cvo}.items()}
    refs["psf"] = (Po.flatten(), N, N)
    for k, (v, nr, nc) in refs.items():
        fits.write_image(os.path.join(chk, k + ".fits"), v, nr, nc)
    exe = os.path.join(ROOT, "moao_amd", "_lib", "check_moao")
    out = os.path.join(d, "out")
    os.makedirs(out)
    args = [exe, "--sys_path", d, "--atm_path", d, "--sys_check", chk, "--compare_all", "1",
            "--write_all", "1", "--out_path", out]
    r = subprocess.run(args + ["--accuracy", "1e300"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = {ln.split(":")[0].strip(): [float(x) for x in ln.split(":")[1].split()]
             for ln in r.stdout.splitlines() if re.match(r"^(F norm|max err)\s*:", ln)}
    names = ["Cmm", "Ctm", "Ctt", "R", "Cee", "Cvv", "psf"]
    for k, (mx, fr) in zip(names, zip(lines["max err"], lines["F norm"])):
        got, _ = fits.read_image(os.path.join(out, k + ".fits"))
        v, nr, nc = refs[k]
        omx, ofr = compare_matrices(got, v, nr, nc, "N", "T")
        print(f"    {k}: max {mx:.3e} (oracle metric {omx:.3e}), F {fr:.3e} ({ofr:.3e})")
        # the table prints 7 significant digits
        assert abs(mx - omx) <= 2e-6 * omx + 1e-300
        assert abs(fr - ofr) <= 2e-6 * ofr + 1e-300
    # the stages themselves: Cmm lower part and Ctt at the matcov tolerance
    assert lines["F norm"][0] <= 1e-12 * np.linalg.norm(Co)
    # a wrong reference PSF fails the gate
    fits.write_image(os.path.join(chk, "psf.fits"), 2.0 * refs["psf"][0], N, N)
    r = subprocess.run(args + ["--accuracy", "0.1"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1 and "bad accuracy" in r.stderr


@pytest.mark.gpu
def test_long_exposure_psf(tmp_path):
    """maxobs = 2: the long-exposure PSF written per target is the mean of the two
    short-exposure PSFs of the run (the reference's TODO at chameleon/moao_chameleon.cpp:
    324-345), bit for bit."""
    from moao_amd import fits
    d = str(tmp_path)
    *_, N = _psf_system(d, nprof=3)
    r = subprocess.run([EXE, "--sys_path", d, "--atm_path", d, "--maxrefine", "1", "--maxobs",
                        "2", "--psf", "1", "--out", d], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    p0, _ = fits.read_image(os.path.join(d, "psf_oloop0_iloop0_gal0.fits"))
    p1, _ = fits.read_image(os.path.join(d, "psf_oloop0_iloop1_gal0.fits"))
    lp_, dims = fits.read_image(os.path.join(d, "psf_nbrefine1_maxobs2_gal0.fits"))
    assert dims == (N, N)
    assert not np.array_equal(p0, p1)
    assert np.array_equal(lp_, (p0 + p1) * 0.5)
    # without short-exposure files the PSFs stay on the device: the same long exposure
    d2 = os.path.join(d, "noshort")
    os.makedirs(d2)
    r = subprocess.run([EXE, "--sys_path", d, "--atm_path", d, "--maxrefine", "1", "--maxobs",
                        "2", "--psf", "1", "--out", d2, "--write_short", "0"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert not os.path.exists(os.path.join(d2, "psf_oloop0_iloop0_gal0.fits"))
    lp2, _ = fits.read_image(os.path.join(d2, "psf_nbrefine1_maxobs2_gal0.fits"))
    assert np.array_equal(lp2, lp_)


@pytest.mark.gpu
def test_three_targets_batched_stages(tmp_path):
    """Three truth directions: the driver solves R for all of them at once, computes every
    target's Ctm/Ctt, one batched Cee/Cvv call and one batched PSF call per observation step.
    R of every target, target 0's Cee/Cvv and every target's PSF against the oracle chain on
    the same matrices (R <= 1e-10, Cee/Cvv <= 1e-10, PSF <= 2e-5 of the peak)."""
    import dataclasses
    from moao_amd import fits, psf_inputs, systems
    from moao_amd import matcov as gm
    from moao_amd.tomo import Tomo
    from oracle import intersample as oi
    from oracle import lapack_path as lp
    from oracle.check import rel_fro
    d = str(tmp_path)
    spec = dataclasses.replace(systems.WORKLOADS["small"], ntarget=3, target_field_diam_as=40.0)
    s, atms = systems.make_system(d, spec, nprof=2)
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
    print(f"[cli] 3 targets n={n} m={m} nact={nact} N={N}", flush=True)
    r = subprocess.run([EXE, "--sys_path", d, "--atm_path", d, "--maxrefine", "1", "--maxobs",
                        "1", "--psf", "1", "--out", d], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    Rf, dims = fits.read_image(os.path.join(d, "R.fits"))
    assert dims == (m, 3 * n)
    Rf = Rf.reshape((m, 3 * n), order="F")
    C = gm.cmm(t)
    Rs = []
    for g in range(3):
        t.set_target(g)
        B = gm.ctm(t)
        Ro = B.copy(order="F")
        assert lp.reconstructor(C.copy(order="F"), Ro) == 0
        Rg = Rf[:, g * n:(g + 1) * n]
        assert rel_fro(Rg, Ro) <= 1e-10, g
        Rs.append(np.asfortranarray(Rg))
    t.update_atm(atms[1])
    C = gm.cmm(t)
    for g in range(3):
        t.set_target(g)
        B, T = gm.ctm(t), gm.ctt(t)
        Cee, Cvv = lp.cee_cvv(C.copy(order="F"), T.copy(order="F"), B.copy(order="F"), Rs[g],
                              np.asfortranarray(Dx))
        if g == 0:
            Ceef, _ = fits.read_image(os.path.join(d, "Cee.fits"))
            Cvvf, _ = fits.read_image(os.path.join(d, "Cvv.fits"))
            assert rel_fro(Ceef.reshape((m, m), order="F"), Cee) <= 1e-10
            assert rel_fro(Cvvf.reshape((nact, nact), order="F"), Cvv) <= 1e-10
        psf, pd = fits.read_image(os.path.join(d, f"psf_oloop0_iloop0_gal{g}.fits"))
        ref = oi.process(Cvv.flatten(order="F"), idx, a2f, otf, N, dm.ndiam)
        err = np.max(np.abs(psf.reshape(N, N) - ref)) / np.max(np.abs(ref))
        print(f"    target {g}: PSF err {err:.2e}", flush=True)
        assert err <= 2e-5, g
