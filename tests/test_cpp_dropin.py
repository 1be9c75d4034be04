"""The header-only MOAO_HIP:: layer compiles against a Tomo_struct-shaped caller and links
against libmoao_hip.so with plain g++ (CPU: link check only; GPU: the refine step through it
gives the same R, bit for bit, as the C-ABI path and matches the oracle's LAPACK path)."""
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
