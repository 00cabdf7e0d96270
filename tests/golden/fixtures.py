"""Readers for the reference fixture files in tests/golden/reference_data.

The formats follow the reference's test loader (cpp/tests/data.in.h:14-55): comma-separated reals,
and complex values written as "(re,im)" pairs.  FITS images are read the way pfitsio::read3d
(cpp/purify/pfitsio.cc:173-210) hands them to the operator: the flat pixel stream in file order
(NAXIS1 fastest) with a zero imaginary part.
"""
import json
import os
import re

import numpy as np

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_data")


def path(*parts):
    return os.path.join(ROOT, *parts)


def read_reals(*parts):
    with open(path(*parts)) as f:
        txt = f.read()
    return np.array([float(t) for t in txt.replace("\n", " ").split(",") if t.strip()], dtype=np.float64)


def read_complex(*parts):
    with open(path(*parts)) as f:
        txt = f.read()
    pairs = re.findall(r"\(\s*([^,()]+)\s*,\s*([^()]+)\s*\)", txt)
    return np.array([complex(float(a), float(b)) for a, b in pairs], dtype=np.complex128)


def read_fits_image(*parts):
    """Minimal FITS primary-HDU reader (BITPIX -32/-64/16/32, big endian)."""
    with open(path(*parts), "rb") as f:
        raw = f.read()
    hdr = {}
    off = 0
    while True:
        block = raw[off: off + 2880]
        off += 2880
        done = False
        for i in range(0, 2880, 80):
            card = block[i: i + 80].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] == "= ":
                val = card[10:].split("/")[0].strip()
                hdr[key] = val
        if done:
            break
    bitpix = int(hdr["BITPIX"])
    naxis = int(hdr["NAXIS"])
    dims = [int(hdr[f"NAXIS{k}"]) for k in range(1, naxis + 1)]
    n = int(np.prod(dims))
    dt = {-32: ">f4", -64: ">f8", 16: ">i2", 32: ">i4"}[bitpix]
    data = np.frombuffer(raw, dtype=dt, count=n, offset=off).astype(np.float64)
    bscale = float(hdr.get("BSCALE", "1").replace("D", "E"))
    bzero = float(hdr.get("BZERO", "0").replace("D", "E"))
    data = data * bscale + bzero
    return data, dims


def convolution_known_answers():
    with open(path("convolution_known_answers.json")) as f:
        return json.load(f)
