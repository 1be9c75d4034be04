"""GPU parity of the matcov generator against the oracle (oracle/matcov_oracle.c, which is
bit-identical to the hand-built reference on these fixtures).

Tolerance: the device pow/exp/sqrt may differ from the host libm in the last ulp; the 4-point
second differences of the structure function amplify that by |D|/|cov| (<= ~1e4 here), so
elements must agree to 1e-12 of the matrix's largest entry and the relative Frobenius error
must be <= 3e-13 (measured: 1.0e-13 on the nx456 2-layer Ctm).
"""
import numpy as np
import pytest

from conftest import FIXTURES, fixture_system
from moao_amd import matcov as gm
from moao_amd.tomo import Tomo
from oracle import matcov as om
from oracle.tomo import Tomo as OTomo

pytestmark = pytest.mark.gpu


def _pair(name):
    s, a = fixture_system(name)
    print(f"[matcov] system {name}", flush=True)
    return Tomo(s, a), OTomo.from_files(s, a)


def _close(g, o, rtol_el=1e-12, rtol_fro=3e-13):
    scale = np.max(np.abs(o))
    el = np.max(np.abs(g - o)) / scale
    fro = np.linalg.norm(g - o) / np.linalg.norm(o)
    assert el <= rtol_el, f"elementwise {el:.3e}"
    assert fro <= rtol_fro, f"frobenius {fro:.3e}"


@pytest.mark.parametrize("name", FIXTURES)
def test_cmm_ctm_ctt_match_oracle(name):
    t, o = _pair(name)
    n, m = t.nmeas(), t.nmeasts()
    # Cmm: lower triangle generated, strictly-upper part left as the caller had it
    sentinel = 7.25
    g = np.full((n, n), sentinel, order="F")
    gm.matcov_tile(g, n, n, 0, 0, n, t, gm.CMM)
    ref = np.full((n, n), sentinel, order="F")
    om.matcov_tile(ref, n, n, 0, 0, n, o, 1)
    assert np.all(g[np.triu_indices(n, 1)] == sentinel)
    _close(np.tril(g), np.tril(ref))
    t.set_target(0)
    o.set_target(0)
    _close(gm.ctm(t), om.ctm(o))
    _close(gm.ctt(t), om.ctt(o))


def test_offset_tiles_and_lda():
    t, o = _pair(FIXTURES[1])
    for (nr, nc, xo, yo, lda, typ) in [(100, 70, 37, 11, 128, 1), (64, 64, 300, 300, 64, 1),
                                       (33, 91, 5, 700, 40, 3), (152, 152, 0, 0, 152, 4),
                                       (17, 19, 100, 3, 17, 4)]:
        if typ == 3:
            t.set_target(0)
            o.set_target(0)
        buf = np.full(lda * nc, -1.0)
        gm.matcov_tile(buf, nr, nc, xo, yo, lda, t, typ)
        ref = np.full(lda * nc, -1.0)
        om.matcov_tile(ref, nr, nc, xo, yo, lda, o, typ)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(buf - ref)) <= 1e-12 * scale, (nr, nc, xo, yo, lda, typ)
        # entries outside the tile rows (lda padding) are untouched
        pad = buf.reshape(nc, lda)[:, nr:]
        assert np.all(pad == -1.0)


def test_bad_type_is_an_error():
    from moao_amd.lib import MoaoError
    t, _ = _pair(FIXTURES[0])
    buf = np.zeros((4, 4), order="F")
    with pytest.raises(MoaoError):
        gm.matcov_tile(buf, 4, 4, 0, 0, 4, t, 5)


def test_cmm_is_spd_and_symmetric_blocks():
    t, _ = _pair(FIXTURES[0])
    c = gm.cmm(t)
    full = np.tril(c) + np.tril(c, -1).T
    w = np.linalg.eigvalsh(full)
    assert w.min() > 0


def test_elt_workload_tiles(tmp_path):
    """The n = 18,688 system (40x40 subapertures, 6 LGS + 2 NGS, 10 layers: elongated LGS next to
    NGS, layers some WFS do not see): 256 x 256 tiles of Cmm (on and off the diagonal, the
    last rows and columns), Ctm and Ctt against the oracle.  As for the whole fixture matrices,
    elements must agree to 1e-12 of the matrix's largest entry (its diagonal: Cmm's own for Cmm,
    the truth sensor's Ctt for Ctm and Ctt).  A per-tile relative norm is not a meaningful bar
    away from the diagonal: the ulp-level device/libm differences of the structure function
    stay at ~1e-14 of the diagonal scale everywhere (measured), while far-apart WFS pairs have
    covariances 1e2-1e5 times smaller than that scale."""
    from moao_amd import systems
    s, a = systems.make_system(str(tmp_path), "elt")
    t, o = Tomo(s, a[0]), OTomo.from_files(s, a[0])
    n, m = t.nmeas(), t.nmeasts()
    print(f"[matcov] elt workload n={n} m={m}", flush=True)
    t.set_target(0)
    o.set_target(0)
    T = 256

    def tile(typ, xo, yo):
        buf = np.full((T, T), -1.0, order="F")
        gm.matcov_tile(buf, T, T, xo, yo, T, t, typ)
        ref = np.full((T, T), -1.0, order="F")
        om.matcov_tile(ref, T, T, xo, yo, T, o, typ)
        if typ == 1:  # only the lower part (row >= col) is generated
            keep = (np.arange(T)[:, None] + xo) >= (np.arange(T)[None, :] + yo)
            assert np.all(buf[~keep] == -1.0)
            buf, ref = np.where(keep, buf, 0.0), np.where(keep, ref, 0.0)
        return buf, ref

    scale = {1: np.max(np.abs(tile(1, 0, 0)[1]))}
    scale[3] = scale[4] = np.max(np.abs(tile(4, 0, 0)[1]))
    tiles = [(1, 0, 0), (1, 9000, 4000), (1, n - T, n - T), (1, 12000, 11900), (1, n - T, 0),
             (1, 3000, 2000), (3, 1000, 15000), (3, m - T, n - T), (4, 0, 0), (4, 2000, 1000)]
    for typ, xo, yo in tiles:
        g, r = tile(typ, xo, yo)
        err = np.max(np.abs(g - r)) / scale[typ]
        print(f"    type {typ} tile at ({xo}, {yo}): {err:.2e} of the diagonal scale", flush=True)
        assert err <= 1e-12


def test_whole_matrix_pair_kernel_is_bit_identical_to_element_tiles(tmp_path):
    """A whole Cmm or Ctm runs the pair kernel (one lane per subaperture pair, shared structure
    functions); any other tile runs the element kernel.  Both must give the same bits: the
    fixtures (LGS only, equal subaperture sizes per layer) and the n = 18,688 system (LGS and
    NGS: unequal sizes on the layers the LGS cone shrinks), column blocks of the whole matrix
    against the same columns generated as tiles at a column offset."""
    from moao_amd import systems
    s, a = systems.make_system(str(tmp_path), "elt")
    cases = [Tomo(*fixture_system(FIXTURES[0])), Tomo(*fixture_system(FIXTURES[1])), Tomo(s, a[0])]
    for t in cases:
        n, m = t.nmeas(), t.nmeasts()
        t.set_target(0)
        print(f"[matcov-pairs] n={n} m={m}", flush=True)
        C = gm.cmm(t)
        B = gm.ctm(t)
        W = 300
        for c0 in sorted({0, n // 3, n - W}):
            blk = np.full((n, W), -1.0, order="F")
            gm.matcov_tile(blk, n, W, 0, c0, n, t, gm.CMM)
            lower = np.arange(n)[:, None] >= np.arange(c0, c0 + W)[None, :]
            assert np.array_equal(blk[lower], C[:, c0:c0 + W][lower]), (n, "Cmm", c0)
            bb = np.empty((m, W), order="F")
            gm.matcov_tile(bb, m, W, 0, c0, m, t, gm.CTM)
            assert np.array_equal(bb, B[:, c0:c0 + W]), (n, "Ctm", c0)
