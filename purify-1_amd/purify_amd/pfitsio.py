"""FITS image I/O (cpp/purify/pfitsio.h / pfitsio.cc: read2d, read3d, write2d, write3d).

The reference goes through cfitsio, which is not in this image; this module reads and writes the
primary HDU itself (FITS 4.0: 2880-byte blocks of 80-character cards, big-endian data).  Images keep
the reference's flat order: the data in the file are the image vector the operators take
(pfitsio.cc:97-100 writes Vector::Map(image.data()), 197-205 reads it back), NAXIS1 fastest.  Here a
plane is returned as a complex128 array of shape (NAXIS2, NAXIS1) in C order, so ``a.ravel()`` is that
vector, and ``write2d(a, ...)`` writes ``a.ravel()`` with NAXIS1 = a.shape[1].
"""
from __future__ import annotations

import datetime
import math
import os
from dataclasses import dataclass

import numpy as np

_BLOCK = 2880
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


@dataclass
class header_params:  # noqa: N801 -- pfitsio.h:17-38
    fits_name: str = "output_image.fits"
    pix_units: str = "Jy/Beam"
    ra: float = 0.0               # radians
    dec: float = 0.0              # radians
    mean_frequency: float = 0.0   # MHz
    cell_x: float = 1.0           # arcsec
    cell_y: float = 1.0           # arcsec
    residual_convergence: float = 1e-4
    relative_variation: float = 1e-3
    polarisation: int = 1         # stokes::I
    channels_total: int = 1
    channel_width: float = 8.0    # MHz
    niters: int = 0
    hasconverged: bool = False
    epsilon: float = 0.0


def _card(key, value=None, comment=""):
    if value is None:
        return f"{key:<80}"[:80]
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, (int, np.integer)):
        v = f"{int(value):>20}"
    elif isinstance(value, (float, np.floating)):
        v = f"{repr(float(value)).upper():>20}"
        if len(v) > 20:
            v = f"{float(value):>20.13E}"
    else:
        s = str(value).replace("'", "''")
        v = f"'{s:<8}'"
    c = f"{key:<8}= {v}" + (f" / {comment}" if comment else "")
    return f"{c:<80}"[:80]


def _parse_header(f):
    cards = {}
    order = []
    while True:
        blk = f.read(_BLOCK)
        if len(blk) < _BLOCK:
            raise RuntimeError("Error opening image: truncated FITS header")
        text = blk.decode("ascii", errors="replace")
        for i in range(0, _BLOCK, 80):
            card = text[i:i + 80]
            key = card[:8].strip()
            if key == "END":
                return cards, order
            if card[8:10] != "= ":
                continue
            val = card[10:]
            if val.lstrip().startswith("'"):
                s = val.lstrip()[1:]
                out, j = [], 0
                while j < len(s):
                    if s[j] == "'":
                        if j + 1 < len(s) and s[j + 1] == "'":
                            out.append("'")
                            j += 2
                            continue
                        break
                    out.append(s[j])
                    j += 1
                cards[key] = "".join(out).rstrip()
            else:
                v = val.split("/")[0].strip()
                if v in ("T", "F"):
                    cards[key] = v == "T"
                else:
                    try:
                        cards[key] = int(v)
                    except ValueError:
                        cards[key] = float(v.replace("D", "E"))
            order.append(key)


def read3d(fits_name):
    """The planes of the primary image (pfitsio.cc:173-211): a list of complex128 arrays (NAXIS2, NAXIS1),
    real part from the file; more than one polarisation is an error, as in the reference."""
    with open(fits_name, "rb") as f:
        h, _ = _parse_header(f)
        naxis = int(h.get("NAXIS", 0))
        if naxis < 1:
            raise RuntimeError("Image contains zero axes.")
        rows = int(h["NAXIS1"])
        cols = int(h["NAXIS2"]) if naxis > 1 else 1
        chans = int(h["NAXIS3"]) if naxis > 2 else 1
        pols = int(h["NAXIS4"]) if naxis > 3 else 1
        if pols > 1:
            raise RuntimeError("Too many polarisations when reading " + fits_name)
        bitpix = int(h["BITPIX"])
        if bitpix not in _DTYPES:
            raise RuntimeError(f"unsupported BITPIX {bitpix} in {fits_name}")
        n = rows * cols * chans
        data = np.frombuffer(f.read(n * abs(bitpix) // 8), dtype=_DTYPES[bitpix], count=n).astype(np.float64)
    data = data * float(h.get("BSCALE", 1.0)) + float(h.get("BZERO", 0.0))
    return [data[i * rows * cols:(i + 1) * rows * cols].reshape(cols, rows).astype(np.complex128)
            for i in range(chans)]


def read2d(fits_name):
    """The first plane (pfitsio.cc:67-76)."""
    return read3d(fits_name)[0]


def read_header(fits_name):
    """The primary header's keyword values."""
    with open(fits_name, "rb") as f:
        return _parse_header(f)[0]


def write3d(images, header, pix_units=None, overwrite=True):
    """pfitsio.cc:77-150: DOUBLE_IMG, NAXIS = 4 (NAXIS1 = a.shape[1], NAXIS2 = a.shape[0], NAXIS3 = planes,
    NAXIS4 = 1), the reference's WCS keywords and PURIFY history.  ``header`` is a header_params or a file
    name (then ``pix_units`` applies, as in the reference's second overload)."""
    if isinstance(header, (str, os.PathLike)):
        header = header_params(fits_name=str(header), pix_units=pix_units or "Jy/Beam")
    planes = [np.real(np.asarray(a)) for a in images]
    a0 = planes[0]
    if a0.ndim == 1:
        raise ValueError("an image is a 2-D array (NAXIS2, NAXIS1)")
    ny, nx = a0.shape
    if any(p.shape != a0.shape for p in planes):
        raise ValueError("every plane must have the same shape")
    if os.path.exists(header.fits_name):
        if not overwrite:
            raise RuntimeError("Problem creating fits file:" + header.fits_name)
        os.remove(header.fits_name)
    cards = [_card("SIMPLE", True), _card("BITPIX", -64), _card("NAXIS", 4), _card("NAXIS1", nx), _card("NAXIS2", ny),
             _card("NAXIS3", len(planes)), _card("NAXIS4", 1), _card("EXTEND", True),
             _card("BUNIT", header.pix_units),
             _card("CRPIX1", int(math.floor(nx / 2)) + 1), _card("CRPIX2", int(math.floor(ny / 2)) + 1),
             _card("CRPIX3", 1), _card("CRPIX4", 1),
             _card("CRVAL1", float(header.ra * 180.0 / math.pi)), _card("CRVAL2", float(header.dec * 180.0 / math.pi)),
             _card("CRVAL3", float(header.mean_frequency * 1e6)), _card("CRVAL4", float(header.polarisation)),
             _card("CTYPE1", "RA--SIN"), _card("CTYPE2", "DEC--SIN"), _card("CTYPE3", "FREQ-OBS"),
             _card("CTYPE4", "STOKES"),
             _card("CDELT1", float(-header.cell_x / 3600.0)), _card("CDELT2", float(header.cell_y / 3600.0)),
             _card("CDELT3", float(header.channel_width * 1e6)), _card("CDELT4", 1.0),
             _card("BTYPE", "intensity"), _card("EQUINOX", 2000),
             f"{'HISTORY SOFTWARE: Purify':<80}",
             f"{'HISTORY PURIFY-NITERS: ' + str(int(header.niters)):<80}",
             f"{'HISTORY PURIFY-CONVERGED: ' + ('True' if header.hasconverged else 'False'):<80}",
             f"{'HISTORY PURIFY-RELATIVEVARIATION: ' + repr(float(header.relative_variation)):<80}",
             f"{'HISTORY PURIFY-RESIDUALCONVERGENCE: ' + repr(float(header.residual_convergence)):<80}",
             f"{'HISTORY PURIFY-EPSILON: ' + repr(float(header.epsilon)):<80}",
             _card("DATE", datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%S")),
             f"{'END':<80}"]
    hdr = "".join(c[:80] for c in cards).encode("ascii")
    hdr += b" " * (-len(hdr) % _BLOCK)
    data = np.concatenate([p.ravel() for p in planes]).astype(">f8").tobytes()
    data += b"\0" * (-len(data) % _BLOCK)
    with open(header.fits_name, "wb") as f:
        f.write(hdr)
        f.write(data)


def write2d(image, header, pix_units=None, overwrite=True):
    """pfitsio.cc:34-65."""
    write3d([image], header, pix_units, overwrite)
