"""Pins the oracle's G, G^H, Z, F, full Phi, SpMV, units and power method.

Reference tests restated:
  cpp/tests/operators.cc:65-92      gridding G / G^H against Matlab fixtures, isApprox 1e-5
  cpp/tests/operators.cc:114-131    FFT round trip, 1e-4
  cpp/tests/operators.cc:147-154    power method on the normalised operator = 1 +- tol
  cpp/tests/measurement_operator.cc:14-84  full Phi on M31, max-normalised, 1e-6
  cpp/tests/utils.cc:590-616        SpMV against a reference product, rel 1e-13
  cpp/tests/wproj_utils.cc:85-125   uv units
Other pins are textbook: a brute-force DFT matrix, Parseval, adjointness, dense SVD.
"""
import numpy as np
import pytest
import scipy.stats

import oracle
from golden import fixtures


def isapprox(a, b, prec):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


def _operators_fixture():
    u = fixtures.read_reals("operators", "u_data")
    v = fixtures.read_reals("operators", "v_data")
    return u, v


def test_gridding_matrix_fixture():
    u, v = _operators_fixture()
    M = len(u)
    assert M == 10
    G = oracle.gridding_matrix(u, v, np.ones(M, np.complex128), 16, 16, 2.0, oracle.KB, 4, 4, 4, 4)
    x = fixtures.read_complex("operators", "direct_input_data")
    assert len(x) == 32 * 32
    y = G.dot(x)
    assert isapprox(y, fixtures.read_complex("operators", "expected_direct_data"), 1e-5)
    GH = G.adjoint()
    yin = fixtures.read_complex("operators", "indirect_input_data")
    g = GH.dot(yin)
    assert isapprox(g, fixtures.read_complex("operators", "expected_indirect_data"), 1e-5)


def test_gridding_adjointness_and_nnz():
    rng = np.random.default_rng(1)
    M, n = 57, 24
    u = rng.uniform(-30, 30, M)
    v = rng.uniform(-30, 30, M)
    wts = rng.normal(size=M) + 1j * rng.normal(size=M)
    G = oracle.gridding_matrix(u, v, wts, n, n, 2.0, oracle.KB, 6, 6, 6, 6)
    assert G.nnz == M * 36
    x = rng.normal(size=(2 * n) ** 2) + 1j * rng.normal(size=(2 * n) ** 2)
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    lhs = np.vdot(y, G.dot(x))
    rhs = np.vdot(G.adjoint().dot(y), x)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_literal_phase_differs_by_roundoff_only():
    rng = np.random.default_rng(2)
    M, n = 200, 256
    u = rng.normal(0, 80, M)
    v = rng.normal(0, 80, M)
    wts = np.ones(M, np.complex128)
    a = oracle.gridding_matrix(u, v, wts, n, n, 2.0, oracle.KB, 4, 4, 4, 4, literal_phase=False)
    b = oracle.gridding_matrix(u, v, wts, n, n, 2.0, oracle.KB, 4, 4, 4, 4, literal_phase=True)
    assert np.array_equal(a.colidx, b.colidx)
    assert np.max(np.abs(a.vals - b.vals)) < 1e-12


@pytest.mark.parametrize("Nv,Nu", [(8, 8), (16, 32), (12, 10), (64, 64)])
def test_fft_matches_dft_matrix(Nv, Nu):
    rng = np.random.default_rng(Nv * 100 + Nu)
    x = rng.normal(size=Nv * Nu) + 1j * rng.normal(size=Nv * Nu)
    Fv = np.exp(-2j * np.pi * np.outer(np.arange(Nv), np.arange(Nv)) / Nv)
    Fu = np.exp(-2j * np.pi * np.outer(np.arange(Nu), np.arange(Nu)) / Nu)
    ref = (Fv @ x.reshape(Nv, Nu) @ Fu.T).ravel() / np.sqrt(Nv * Nu)
    got = oracle.fft2d(Nv, Nu, x, -1)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)
    back = oracle.fft2d(Nv, Nu, got, +1)
    assert np.linalg.norm(back - x) <= 1e-13 * np.linalg.norm(x)


def test_fft_plane_wave_and_parseval():
    N = 128
    k0, l0 = 5, 17
    p, q = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    x = np.exp(2j * np.pi * (k0 * p + l0 * q) / N).ravel()
    X = oracle.fft2d(N, N, x, -1).reshape(N, N)
    expect = np.zeros((N, N), np.complex128)
    expect[k0, l0] = N  # unitary: N^2 / sqrt(N^2)
    assert np.max(np.abs(X - expect)) < 1e-10
    rng = np.random.default_rng(3)
    y = rng.normal(size=N * N) + 1j * rng.normal(size=N * N)
    assert np.linalg.norm(oracle.fft2d(N, N, y)) == pytest.approx(np.linalg.norm(y), rel=1e-13)


def test_fft_round_trip_reference_tolerance():
    # cpp/tests/operators.cc:114-131 (16x16 image, oversample 2)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, 1024) + 1j * rng.uniform(-1, 1, 1024)
    assert isapprox(oracle.fft2d(32, 32, oracle.fft2d(32, 32, x, -1), +1), x, 1e-4)


def test_zero_pad_adjoint():
    rng = np.random.default_rng(5)
    ny, nx = 10, 14
    S = oracle.correction2d(2.0, ny, nx, oracle.KB, 4, 4)
    x = rng.normal(size=ny * nx) + 1j * rng.normal(size=ny * nx)
    g = rng.normal(size=4 * ny * nx) + 1j * rng.normal(size=4 * ny * nx)
    assert abs(np.vdot(g, oracle.zero_pad(S, ny, nx, 2.0, x)) - np.vdot(oracle.zero_pad_adjoint(S, ny, nx, 2.0, g), x)) \
        < 1e-12 * np.linalg.norm(g) * np.linalg.norm(x) * np.max(S)
    z = oracle.zero_pad(S, ny, nx, 2.0, x).reshape(2 * ny, 2 * nx)
    assert np.count_nonzero(z) == ny * nx
    np.testing.assert_allclose(z[ny // 2: ny // 2 + ny, nx // 2: nx // 2 + nx].ravel(), S * x, rtol=1e-15)


@pytest.mark.parametrize("kernel,J", [("kb", 4), ("pswf", 6), ("gauss", 6)])
def test_full_operator_m31_fixture(kernel, J):
    # cpp/tests/measurement_operator.cc: the test passes (v, u) into the (u, v) slots
    u = fixtures.read_reals("measurement_operator", "u")
    v = fixtures.read_reals("measurement_operator", "v")
    img, dims = fixtures.read_fits_image("images", "M31.fits")
    assert dims[:2] == [256, 256]
    x = img.astype(np.complex128)
    M = len(u)
    op = oracle.Operator(v, u, np.zeros(M), np.ones(M, np.complex128), 256, 256, 2.0,
                         oracle.KERNELS[kernel], J, J, power_iters=1, power_tol=1e-9)
    y = op.forward(x)
    yexp = fixtures.read_complex("measurement_operator", f"{kernel}_exact_data")
    assert len(yexp) == M
    assert isapprox(y / np.max(np.abs(y)), yexp / np.max(np.abs(yexp)), 1e-6)


def _dense_phi(op, n):
    cols = []
    for k in range(n * n):
        e = np.zeros(n * n, np.complex128)
        e[k] = 1
        cols.append(op.forward(e))
    return np.array(cols).T


def test_operator_adjoint_and_power_method_vs_svd():
    rng = np.random.default_rng(6)
    n, M = 8, 40
    u = rng.uniform(-8, 8, M)
    v = rng.uniform(-8, 8, M)
    wts = rng.normal(size=M) + 1j * rng.normal(size=M)
    op = oracle.Operator(u, v, None, wts, n, n, 2.0, oracle.KB, 4, 4)
    A = _dense_phi(op, n)
    smax = np.linalg.svd(A, compute_uv=False)[0]
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    np.testing.assert_allclose(op.adjoint(y), A.conj().T @ y, rtol=1e-12, atol=1e-14 * np.linalg.norm(y))
    nrm = op.power_method(2000, 1e-9)
    assert nrm == pytest.approx(smax, rel=1e-8)
    op.set_norm(nrm)
    # operators.cc:147-154: the normalised operator has norm 1 +- tol
    assert abs(op.power_method(2000, 1e-9, rng.normal(size=n * n) + 0j, normalised=True) - 1.0) < 1e-8


def test_power_method_zero_iterations_returns_one():
    op = oracle.Operator(np.array([0.3]), np.array([0.2]), None, None, 4, 4)
    assert op.power_method(0, 1e-4) == 1.0


def test_spmv_against_dense():
    rng = np.random.default_rng(7)
    rows, cols, nz = 300, 5000, 4000
    r = rng.integers(0, rows, nz)
    c = rng.integers(0, cols, nz)
    val = rng.normal(size=nz) + 1j * rng.normal(size=nz)
    D = np.zeros((rows, cols), np.complex128)
    np.add.at(D, (r, c), val)
    # CSR from the dense matrix
    rp = [0]
    ci, vv = [], []
    for i in range(rows):
        nzc = np.nonzero(D[i])[0]
        ci += list(nzc)
        vv += list(D[i, nzc])
        rp.append(len(ci))
    A = oracle.CSR(rows, cols, np.array(rp, np.int64), np.array(ci, np.int32), np.array(vv))
    x = rng.normal(size=cols) + 1j * rng.normal(size=cols)
    ref = D @ x
    got = A.dot(x)
    nzr = np.abs(ref) > 0
    assert np.max(np.abs((got - ref)[nzr] / ref[nzr])) < 1e-13


@pytest.mark.parametrize("cell", [1.0, 3.0])
def test_uv_units(cell):
    # cpp/tests/wproj_utils.cc:85-125: u_lambda == u_pixels * pixel_to_lambda(cell, n, os)
    n, os_ = 128, 2.0
    u = np.ones(5)
    up, vp = oracle.uv_to_pixels(u, u, 0, cell, cell, n, n, os_)
    scale = 60.0 * 60.0 * 180.0 / cell / np.floor(os_ * n) / np.pi
    assert isapprox(u, up * scale, 1e-6)


def test_uv_units_radians_take_uv_scale_only():
    # operators.h:408-410: radians go through uv_scale alone, u / (2 pi) * floor(os n), whatever the cell
    n, os_ = 100, 2.0
    u = np.array([-np.pi, -1.0, 0.0, 0.5, np.pi])
    for cell in (1.0, 7.0):
        up, vp = oracle.uv_to_pixels(u, -u, 1, cell, cell, n, n, os_)
        assert isapprox(up, u / (2 * np.pi) * 200, 1e-15) and isapprox(vp, -u / (2 * np.pi) * 200, 1e-15)
    # pixels pass through unchanged
    up, _ = oracle.uv_to_pixels(u, u, 2, 1.0, 1.0, n, n, os_)
    assert np.array_equal(up, u)


def test_l2_radius_matches_chi_distribution():
    # utilities.cc:396-421 uses an asymptotic series for the chi mean; compare to the exact
    # chi(2M) distribution's mean + 2 sd.
    for M in (100, 10_000, 1_000_000):
        exact = scipy.stats.chi(2 * M).mean() + 2 * scipy.stats.chi(2 * M).std()
        assert oracle.calculate_l2_radius(M, 0.0, 2.0) == pytest.approx(exact, rel=1e-6)
    assert oracle.calculate_l2_radius(50, 2.0, 2.0, "chi^2") == pytest.approx(np.sqrt(100 + 2 * np.sqrt(200)) * 2)


def test_snr_to_standard_deviation():
    y = np.full(400, 3.0 + 4.0j)
    assert oracle.snr_to_standard_deviation(y, 30.0) == pytest.approx(5.0 * 10 ** (-1.5), rel=1e-14)
