"""Shared helpers of the GPU parity tests: build a GPU operator and its oracle twin on the same inputs."""
import numpy as np

import oracle
from purify_amd import measurementoperator, utilities

KIND = {"kb": oracle.KB, "gauss": oracle.GAUSS, "box": oracle.BOX, "pswf": oracle.PSWF, "kbmin": oracle.KBMIN,
        "gauss_alt": oracle.GAUSS_ALT}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def isapprox(a, b, prec):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


def make_uv(seed, M, n, sigma=np.pi / 3, os_=2.0):
    u, v = utilities.synth_uv(seed, M, sigma)
    return utilities.radians_to_pixels(u, n, os_), utilities.radians_to_pixels(v, n, os_)


def pair(u, v, ny, nx, weights=None, kernel="kb", Ju=4, Jv=4, os_=2.0, precision="c128", tile=(0, 0)):
    """GPU operator (no normalisation) and the oracle operator on the same pixel coordinates."""
    gpu = measurementoperator.init_degrid_operator_2d(u, v, None, weights, ny, nx, os_, 0, 1e-4, kernel, Ju, Jv,
                                                      precision=precision, tile=tile)
    orc = oracle.Operator(u, v, None, weights, ny, nx, os_, KIND[kernel], Ju, Jv)
    return gpu, orc
