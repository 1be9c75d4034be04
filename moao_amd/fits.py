"""Minimal FITS image I/O in the layout the reference's tools read and write.

The reference stores every matrix as a primary-HDU 2-D image of doubles through cfitsio
(common/io/fits.cpp:110-148 writeFits: BITPIX = DOUBLE_IMG, NAXIS1 = nx, NAXIS2 = ny, the
buffer written in memory order), and reads idx/abs2fi/otftel/Dx the same way
(common/tomo_struct/intersample.cpp:236-300, fits.cpp:66-105).  So a column-major m x n matrix
written as (nx=m, ny=n) keeps its memory order on disk, and the PSF (N x N, `dphi`) likewise.

Format (FITS standard 4.0): 80-character ASCII header cards in 2880-byte blocks ending with
END, then the data big-endian (IEEE double for BITPIX = -64, two's-complement for 8/16/32/64)
padded with zeros to a multiple of 2880 bytes.  Only what those files need is supported: one
primary image, no scaling keywords (BZERO/BSCALE must be absent or identity).
"""
from __future__ import annotations

import numpy as np

BLOCK = 2880
_BITPIX = {np.dtype(np.float64): -64, np.dtype(np.float32): -32, np.dtype(np.int64): 64,
           np.dtype(np.int32): 32, np.dtype(np.int16): 16, np.dtype(np.uint8): 8}
_DTYPE = {-64: ">f8", -32: ">f4", 64: ">i8", 32: ">i4", 16: ">i2", 8: "u1"}


def _card(key: str, value) -> bytes:
    if isinstance(value, bool):
        v = "T" if value else "F"
    else:
        v = str(value)
    c = f"{key:<8}= {v:>20}"
    return c.ljust(80).encode("ascii")


def write_image(path: str, data: np.ndarray, nx: int, ny: int) -> None:
    """Write `data` (nx*ny elements, in its memory order) as an nx x ny FITS image; float
    arrays are written as doubles (what writeFits does for every T)."""
    a = np.asarray(data)
    if a.size != nx * ny:
        raise ValueError(f"{a.size} elements cannot form a {nx} x {ny} image")
    if a.dtype.kind == "f":
        a = a.astype(np.float64)
    flat = np.ravel(a, order="K" if a.flags.f_contiguous or a.flags.c_contiguous else "C")
    bitpix = _BITPIX[np.dtype(flat.dtype)]
    hdr = b"".join([_card("SIMPLE", True), _card("BITPIX", bitpix), _card("NAXIS", 2),
                    _card("NAXIS1", nx), _card("NAXIS2", ny), _card("EXTEND", True),
                    b"END".ljust(80)])
    hdr = hdr.ljust(-(-len(hdr) // BLOCK) * BLOCK, b" ")
    body = flat.astype(_DTYPE[bitpix]).tobytes()
    body = body.ljust(-(-len(body) // BLOCK) * BLOCK, b"\0")
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(body)


def read_image(path: str):
    """Returns (data as a flat array in file order, (NAXIS1, NAXIS2)); doubles for float
    images, int64 for integer images (as the reference reads them with TDOUBLE / TLONG)."""
    with open(path, "rb") as f:
        raw = f.read()
    keys = {}
    pos = 0
    while True:
        if pos + 80 > len(raw):
            raise ValueError(f"{path}: no END card")
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        k = card[:8].strip()
        if k == "END":
            break
        if card[8:10] == "= ":
            keys[k] = card[10:].split("/")[0].strip()
    pos = -(-pos // BLOCK) * BLOCK
    if keys.get("SIMPLE") != "T":
        raise ValueError(f"{path}: not a FITS primary HDU")
    for k, ident in (("BZERO", 0.0), ("BSCALE", 1.0)):
        if k in keys and float(keys[k]) != ident:
            raise ValueError(f"{path}: scaled images are not supported")
    bitpix = int(keys["BITPIX"])
    naxis = int(keys["NAXIS"])
    dims = [int(keys[f"NAXIS{i + 1}"]) for i in range(naxis)]
    n = int(np.prod(dims)) if dims else 0
    dt = np.dtype(_DTYPE[bitpix])
    a = np.frombuffer(raw, dtype=dt, count=n, offset=pos)
    a = a.astype(np.float64) if bitpix < 0 else a.astype(np.int64)
    while len(dims) < 2:
        dims.append(1)
    return a, (dims[0], dims[1])
