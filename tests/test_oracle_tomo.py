"""Pins the oracle's Tomo_struct restatement and the library's C++ host model (CPU only)
against the golden vectors written by the hand-built reference, and the compareMatrices
restatement against the reference's own outputs.

Geometry/projection vectors must match bit for bit.  The LGS noise terms may differ in the last
few ulps (the reference build and Python/C++ here evaluate pow/cos/sin of the same arguments
through different libm call patterns), so they are checked to 4 ulp.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, SYSTEMS
from moao_amd.tomo import Tomo as CTomo
from oracle.check import compare_matrices
from oracle.tomo import Tomo

NAMES = sorted(os.listdir(SYSTEMS))
EXACT = ["X", "Y", "u", "v", "sspSizeL", "cn2", "Nsubap", "Nssp", "lgsExt", "lgsTheta"]
ULP = ["noiseNGS", "noiseLGSxx", "noiseLGSyy", "noiseLGSxy"]
CNAME = {"L0diff": "L0"}


def _files(name):
    d = os.path.join(SYSTEMS, name)
    return os.path.join(d, "sys-params.txt"), os.path.join(d, "prof-1-atmos-night0.txt")


def _cmp(got, ref, name, exact):
    got = np.asarray(got, dtype=float)
    assert got.shape == ref.shape, name
    if exact:
        assert np.array_equal(got, ref), name
    else:
        assert np.all(np.abs(got - ref) <= 4 * np.spacing(np.abs(ref))), name


@pytest.mark.parametrize("name", NAMES)
def test_oracle_tomo_matches_reference(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    t = Tomo.from_files(*_files(name))
    t.set_target(0)
    for f in EXACT + ULP:
        _cmp(getattr(t, f), g["f_" + f], f, f in EXACT)
    nL = t.atm.nLayer
    assert np.array_equal(t.L0diff, g["f_L0diff"][:nL])  # reference vector is oversized
    dims = g["dims"]
    assert (t.nmeas(), t.nmeasts()) == (dims[5], dims[6])


@pytest.mark.parametrize("name", NAMES)
def test_library_tomo_matches_reference(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    t = CTomo(*_files(name))
    t.set_target(0)
    for f in EXACT + ULP:
        _cmp(t.field(CNAME.get(f, f)), g["f_" + f], f, f in EXACT)
    assert np.array_equal(t.field("L0"), g["f_L0diff"][: t.dims()["nLayer"]])
    dims = g["dims"]
    assert (t.nmeas(), t.nmeasts()) == (dims[5], dims[6])


def test_library_tomo_errors(tmp_path):
    from moao_amd.lib import MoaoError
    with pytest.raises(MoaoError):
        CTomo(str(tmp_path / "missing.txt"), str(tmp_path / "missing2.txt"))
    s, a = _files(NAMES[0])
    lines = open(s).read().split("\n")
    trunc = tmp_path / "sys.txt"
    trunc.write_text("\n".join(lines[:30]))
    with pytest.raises(MoaoError):
        CTomo(str(trunc), a)
    t = CTomo(s, a)
    with pytest.raises(MoaoError):
        t.set_target(5)


@pytest.mark.parametrize("name", NAMES)
def test_compare_matrices_matches_reference(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    t = Tomo.from_files(*_files(name))
    from oracle import matcov as om
    ctt = om.ctt(t)
    m = ctt.shape[0]
    k = 0
    for uplo in "NLU":
        for tr in "NT":
            me, fr = compare_matrices(ctt, g["cmp_pert"], m, m, uplo, tr)
            assert me == pytest.approx(g["cmp"][k][0], rel=1e-12)
            assert fr == pytest.approx(g["cmp"][k][1], rel=1e-12)
            k += 1
