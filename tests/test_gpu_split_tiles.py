"""Stored-row gridding with split tiles: on a small grid with many visibilities (a 128^2 image, 256^2 grid,
300000 visibilities around the centre, w-projection rows of 9 x 9 taps), the tiles whose gridding lists
exceed a chunk are cut into parts whose partial sums are added in a fixed order.  Against the oracle on a
20000-visibility subset (the forward rows of the subset; the adjoint of a vector that is zero outside it),
complex128 at 1e-12 and complex64 at 1e-5 (relative L2); the adjoint is the same bit for bit on two runs;
the degridding's chunked launch slots leave the forward rows exactly those of the unchunked sums (checked
against the oracle rows).
"""
import numpy as np
import pytest

import oracle
from gpu_helpers import make_uv, rel
from purify_amd import measurementoperator, utilities

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_split_tiles_match_the_oracle(precision):
    n, M, J, Jw = 128, 300_000, 4, 6
    cell = 3600.0 / n
    u, v = make_uv(1301, M, n, sigma=0.5)
    w = np.random.default_rng(1302).uniform(0.0, 200.0, M)
    lt = measurementoperator.init_degrid_operator_2d(u, v, w, None, n, n, 2.0, 0, 1e-4, "kb", J, J, True, Jw, cell,
                                                     cell, precision=precision)
    assert lt.op.get_tuning("split_tiles") > 0
    x = utilities.synth_image(1303, n, n, 10)
    y = utilities.synth_cnormal(1304, M)
    idx = np.sort(np.random.default_rng(1305).choice(M, size=20000, replace=False))
    one = np.ones(idx.size, np.complex128)
    G = oracle.gridding_matrix_w(u[idx], v[idx], w[idx] - w.mean(), one, n, n, 2.0, oracle.KB, J, J, J, J, Jw,
                                 oracle.W_APPROX, cell, cell)
    S = oracle.correction2d(2.0, n, n, oracle.KB, J, J)
    Nv, Nu = oracle.grid_dims(n, n, 2.0)
    yo = G.dot(oracle.fft2d(Nv, Nu, oracle.zero_pad(S, n, n, 2.0, x), -1))
    assert rel((lt * x)[idx], yo) < TOL[precision]
    ysub = np.zeros(M, np.complex128)
    ysub[idx] = y[idx]
    xo = oracle.zero_pad_adjoint(S, n, n, 2.0, oracle.fft2d(Nv, Nu, G.adjoint().dot(y[idx]), +1))
    xa = lt.adjoint() * ysub
    assert rel(xa, xo) < TOL[precision]
    # the full adjoint, twice: the parts are summed in a fixed order
    a1 = lt.adjoint() * y
    a2 = lt.adjoint() * y
    assert np.array_equal(a1, a2)
    lt.op.close()
