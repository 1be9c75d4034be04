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
