"""GPU chirp-based w-projection matrix (wproj_utilities::wprojection_matrix, cpp/purify/wproj_utilities.cc:132-188)
and operators on a given sparse gridding matrix.

The matrix builder runs in FP64 and is compared with the oracle's dense restatement entry by entry (1e-12
relative per row, identical sparsity) and with the reference's own known answer: at w = 0 the chirp is a
unit delta and the matrix of the identity is the identity (cpp/tests/wproj_utils.cc:187-216).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from gpu_helpers import make_uv, rel
from purify_amd import measurementoperator, utilities, wproj_utilities
from purify_amd._lib import PurifyError

pytestmark = pytest.mark.gpu


def oracle_G(u, v, n, J=4):
    M = len(u)
    G = oracle.gridding_matrix(u, v, np.ones(M, np.complex128), n, n, 2.0, oracle.KB, J, J, J, J)
    return G, sp.csr_matrix((G.vals, G.colidx, G.rowptr), shape=(G.rows, G.cols))


def test_identity_at_w_zero_and_growth_for_random_w():
    # cpp/tests/wproj_utils.cc:187-216: 12 rows of the identity on a 32 x 32 grid, cells of 60 arcsec
    Nx = Ny = 32
    rows = 12
    I = sp.csr_matrix((np.ones(rows, np.complex128), (np.arange(rows), np.arange(rows))), shape=(rows, Nx * Ny))
    w = np.random.default_rng(5).uniform(-1, 1, rows)
    G_id = wproj_utilities.wprojection_matrix(I, Nx, Ny, w * 0, 60, 60, 1, 1)
    G = wproj_utilities.wprojection_matrix(I, Nx, Ny, w, 60, 60, 1, 1)
    assert G_id.nnz == I.nnz
    assert G.nnz >= I.nnz
    coo = G_id.tocoo()
    assert np.all(coo.row == coo.col)
    assert np.max(np.abs(coo.data - 1.0)) < 1e-12


@pytest.mark.parametrize("e_chirp,e_wproj", [(1.0, 1.0), (0.999, 1.0), (0.999, 0.99), (0.95, 0.9)])
def test_matrix_vs_oracle_dense(e_chirp, e_wproj):
    n, M = 16, 48  # 32 x 32 grid
    u, v = make_uv(401, M, n)
    w = np.random.default_rng(402).uniform(0, 200, M)
    cell = 3600.0 / n * 4  # a wide field makes the chirps spread over many cells
    Go, Gs = oracle_G(u, v, n)
    GW = wproj_utilities.wprojection_matrix(Gs, 2 * n, 2 * n, w, cell, cell, e_chirp, e_wproj)
    dense, mask = oracle.wprojection_matrix_dense(Go, 2 * n, 2 * n, w, cell, cell, e_chirp, e_wproj)
    D = GW.toarray()
    # A threshold that coincides with a magnitude to rounding (the bisection's stopping rule can fail and
    # leave it an ulp below the largest of several equal magnitudes of the chirp's symmetric spectrum)
    # makes the keep decisions depend on last-bit rounding; every row without such a tie must match the
    # oracle exactly, in pattern and values.
    checked = 0
    for m in range(M):
        if threshold_margin(Go, m, 2 * n, w[m], cell, e_chirp, e_wproj) < 1e-12:
            continue
        assert np.array_equal(D[m] != 0, mask[m])
        assert rel(D[m], dense[m]) < 1e-12
        checked += 1
    assert checked >= (M if e_chirp >= 0.999 else 4)


def threshold_margin(Go, m, Nside, w, cell, e_chirp, e_wproj):
    """Smallest | |c| / threshold - 1 | over the chirp and the convolved row of visibility m (oracle)."""
    N = Nside * Nside
    ch = oracle.fft2d(Nside, Nside, oracle.generate_chirp(w, cell, cell, Nside, Nside)) * np.sqrt(N)
    a = np.abs(ch)
    keep = max(a.max() * 1e-10, oracle.sparsify_row_dense_thres(ch, e_chirp))
    marg = np.min(np.abs(a / keep - 1)) if keep > 0 else np.inf
    ch = np.where(a > keep, ch, 0)
    lo, hi = Go.rowptr[m], Go.rowptr[m + 1]
    row = oracle.row_wise_convolution(Go.colidx[lo:hi], Go.vals[lo:hi], ch, Nside, Nside)
    b = np.abs(row[row != 0])
    t2 = oracle.sparsify_row_thres(b, e_wproj)
    if t2 > 0:
        marg = min(marg, np.min(np.abs(b / t2 - 1)))
    return marg


def test_matrix_errors():
    G = sp.csr_matrix((4, 32 * 16), dtype=np.complex128)
    with pytest.raises(PurifyError, match="square"):
        wproj_utilities.wprojection_matrix(G, 32, 16, np.zeros(4), 60, 60)
    G = sp.csr_matrix((4, 32 * 32), dtype=np.complex128)
    with pytest.raises(RuntimeError, match="one entry per row"):
        wproj_utilities.wprojection_matrix(G, 32, 32, np.zeros(3), 60, 60)
    with pytest.raises(PurifyError, match="outside the grid"):
        bad = sp.csr_matrix((np.ones(1, np.complex128), (np.array([0]), np.array([32 * 32]))), shape=(4, 32 * 32 + 1))
        wproj_utilities.wprojection_matrix(bad, 32, 32, np.zeros(4), 60, 60)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_operator_on_the_kaiser_bessel_matrix_equals_the_built_operator(precision):
    n, M = 64, 3000
    u, v = make_uv(411, M, n)
    _, Gs = oracle_G(u, v, n, J=4)
    lg = measurementoperator.init_degrid_operator_2d_from_matrix(Gs, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                                 precision=precision)
    orc = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, 4, 4)
    tol = 1e-12 if precision == "c128" else 1e-5
    x = utilities.synth_image(412, n, n, 20)
    assert rel(lg * x, orc.forward(x)) < tol
    y = utilities.synth_cnormal(413, M)
    assert rel(lg.adjoint() * y, orc.adjoint(y)) < tol


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_operator_on_the_chirp_matrix(precision):
    n, M = 16, 200
    u, v = make_uv(421, M, n)
    w = np.random.default_rng(422).uniform(0, 200, M)
    Go, Gs = oracle_G(u, v, n)
    GW = wproj_utilities.wprojection_matrix(Gs, 2 * n, 2 * n, w, 225.0, 225.0, 0.999, 0.99)
    lt = measurementoperator.init_degrid_operator_2d_from_matrix(GW, n, n, 2.0, 200, 1e-7, "kb", 4, 4,
                                                                 precision=precision)
    # Phi = GW F Z S / norm from the oracle's pieces
    S = oracle.correction2d(2.0, n, n, oracle.KB, 4, 4)
    x = utilities.synth_cnormal(423, n * n)
    grid = oracle.fft2d(2 * n, 2 * n, oracle.zero_pad(S, n, n, 2.0, x))
    y_ref = GW.dot(grid) / lt.norm
    tol = 1e-12 if precision == "c128" else 1e-5
    assert rel(lt * x, y_ref) < tol
    yv = utilities.synth_cnormal(424, M)
    lhs = np.vdot(yv, lt * x)
    rhs = np.vdot(lt.adjoint() * yv, x)
    assert abs(lhs - rhs) / abs(lhs) < tol
    # the normalised operator has norm 1 within the power method's tolerance (cpp/tests/operators.cc:147-154)
    assert abs(lt.op.power_method(500, 1e-8) - 1.0) < 1e-4


def test_chirp_rows_apply_the_exact_w_screen_to_a_point_source():
    # Independent of the oracle: with both energy fractions 1 and rows of G holding only the zero-frequency
    # cell, Phi_w delta_b = exp(-2 pi i w (sqrt(1 - X^2 - Y^2) - 1)) Phi delta_b at the direction cosines
    # X = (b - N/2 - 1/2) L / N, L = 2 sin(N cell / 2), the chirp grid of wproj_utilities.cc:37-75 seen
    # through the circular row convolution (the same identity pins the oracle in test_oracle_wproj_pins.py).
    N, cell = 64, 600.0
    w = np.array([500.0, -300.0, 50.0, 1e3])
    M = len(w)
    G = sp.csr_matrix((np.ones(M, np.complex128), (np.arange(M), np.zeros(M, int))), shape=(M, N * N))
    GW = wproj_utilities.wprojection_matrix(G, N, N, w, cell, cell, 1.0, 1.0).toarray()
    L = 2 * np.sin(cell * N * np.pi / (180 * 3600) / 2)
    for bp, bq in ((21, 39), (10, 50), (33, 31), (40, 12), (5, 60)):
        z = np.zeros((N, N), np.complex128)
        z[bp, bq] = 1.0
        g = (np.fft.fft2(z) / N).ravel()  # the unitary forward FFT of the grid
        X, Y = (bq - N / 2 - 0.5) * L / N, (bp - N / 2 - 0.5) * L / N
        want = np.exp(-2j * np.pi * w * (np.sqrt(1 - X * X - Y * Y) - 1)) / N
        np.testing.assert_allclose(GW @ g, want, rtol=1e-12, atol=0)
