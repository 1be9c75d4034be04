"""FITS image I/O in the reference's layout (primary-HDU 2-D images, doubles, memory order):
round trips and the byte layout fixed by the FITS standard (CPU only)."""
import numpy as np

from moao_amd import fits


def test_round_trip_column_major_matrix(tmp_path):
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.standard_normal((37, 53)))
    p = str(tmp_path / "R.fits")
    fits.write_image(p, A, 37, 53)  # nx = rows: memory order kept (writeFits(nx, ny, data))
    a, (nx, ny) = fits.read_image(p)
    assert (nx, ny) == (37, 53)
    assert np.array_equal(a.reshape((37, 53), order="F"), A)


def test_layout_follows_the_standard(tmp_path):
    p = str(tmp_path / "x.fits")
    fits.write_image(p, np.arange(6, dtype=np.float64), 3, 2)
    raw = open(p, "rb").read()
    assert len(raw) == 2 * fits.BLOCK                      # one header + one data block
    cards = [raw[i:i + 80].decode() for i in range(0, fits.BLOCK, 80)]
    assert cards[0].startswith("SIMPLE  =") and cards[0][29] == "T"   # logical in column 30
    assert cards[1].startswith("BITPIX  =") and cards[1][10:30].strip() == "-64"
    assert any(c.startswith("END ") for c in cards)
    data = np.frombuffer(raw[fits.BLOCK:fits.BLOCK + 48], dtype=">f8")
    assert np.array_equal(data, np.arange(6.0))             # big-endian IEEE doubles
    assert raw[fits.BLOCK + 48:] == b"\0" * (fits.BLOCK - 48)


def test_integer_index_image(tmp_path):
    idx = np.arange(1, 101, dtype=np.int64)
    p = str(tmp_path / "idx.fits")
    fits.write_image(p, idx, 100, 1)
    a, dims = fits.read_image(p)
    assert dims == (100, 1) and a.dtype == np.int64 and np.array_equal(a, idx)


def test_command_matrix_is_the_filtered_pseudo_inverse():
    """Dx from the truncated SVD (commande.py:108-127) of the truth sensor's interaction matrix
    (commande.py:81-106): nact x nmeasts; with no mode filtered it is the Moore-Penrose inverse
    (numpy's pinv, and Dx M Dx = Dx, M Dx M = M); filtering by a condition number drops exactly
    the singular values below max/sqrt(condi).  The interaction matrix itself: a pure tip
    (actuator voltages proportional to x) gives equal x slopes on every subaperture away from the
    edge and ~zero y slopes."""
    import numpy as np
    from moao_amd import psf_inputs, systems
    from moao_amd.tomo import Tomo
    import tempfile
    d = tempfile.mkdtemp()
    s, atms = systems.make_system(d, "small")
    t = Tomo(s, atms[0])
    m = t.nmeasts()
    X, Y = t.field("X"), t.field("Y")
    ns = m // 2
    Xt, Yt = X[-ns:], Y[-ns:]
    dm = psf_inputs.make_dm(10)
    M = psf_inputs.interaction_matrix(Xt, Yt, 10, 8.0, dm)
    assert M.shape == (m, dm.nact)
    Dx = psf_inputs.command_matrix(M, thresh=1e-12)
    assert Dx.shape == (dm.nact, m)
    w = np.linalg.svd(M, compute_uv=False)
    assert w.min() > 1e-12
    P = np.linalg.pinv(M)
    assert np.linalg.norm(Dx - P) <= 1e-9 * np.linalg.norm(P)
    assert np.linalg.norm(Dx @ M @ Dx - Dx) <= 1e-9 * np.linalg.norm(Dx)
    assert np.linalg.norm(M @ Dx @ M - M) <= 1e-9 * np.linalg.norm(M)
    condi = (w.max() / w[len(w) // 2]) ** 2 * 1.0001
    Dc = psf_inputs.command_matrix(M, condi=condi)
    kept = int(np.sum(w >= w.max() / np.sqrt(condi)))
    assert np.linalg.matrix_rank(Dc, tol=1e-12 * np.linalg.norm(Dc)) == kept
    cx, _ = psf_inputs.actuator_positions(dm)
    sl = M @ cx
    inner = np.hypot(Xt, Yt) < 0.6 * 4.0
    assert np.std(sl[:ns][inner]) <= 0.05 * abs(np.mean(sl[:ns][inner]))
    assert np.max(np.abs(sl[ns:][inner])) <= 0.05 * abs(np.mean(sl[:ns][inner]))
