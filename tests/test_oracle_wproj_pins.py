"""Pins of the oracle's w != 0 projection by what the mathematics fixes (CPU, no GPU).

The reference's own tests hold no w != 0 known answer (cpp/tests/wproj_utils.cc:41-66 only runs the
projection), so the checks below come from the measurement equation and from properties the algorithm
must have, and each is run against planted mistakes that it must catch.

Approximate w-kernel rows (w_term = true; projection_kernels.cc:16-85, operators.cc:49-99):
  - a w-kernel sample is -1 times the 2-D Fourier transform of the small-angle w-screen
    exp(i pi w (l^2 + m^2)) (the measurement equation's exp(-2 pi i w (n - 1)) with n - 1 ~ -(l^2+m^2)/2),
    evaluated from Fresnel integrals (scipy.special.fresnel), at the pixel offset times the grid's uv cell.
    The reference writes 1/(i w) where the transform has i/w, hence the -1 (projection_kernels.cc:16-22);
  - the grid's uv cell is 1 / (the grid's field of view in radians), and equals the ratio of wavelengths
    to pixels of the operator's own unit conversion (utilities.cc:208-278);
  - outside |u du / w| < 1 and |v dv / w| < 1 the sample is 1 (projection_kernels.cc:37-45);
  - a box w-kernel of Jw taps smears the plain Kaiser-Bessel row over Jw integer shifts centred on the
    visibility (odd Jw: shifts -(Jw-1)/2 .. (Jw-1)/2);
  - the operator sees w - mean(w) (operators.h:323): adding a constant to w changes nothing.
Chirp-based rows (wprojection_matrix, wproj_utilities.cc:76-188, wproj_utilities.h:89-287):
  - with both energy fractions 1 and a row of G that holds the zero-frequency cell only, Phi_w applied
    to a point source equals the plain measurement times the exact w-screen
    exp(-2 pi i w (sqrt(1 - X^2 - Y^2) - 1)) at the source's direction cosines.  The row convolution is a
    circular convolution of G's row with the chirp's DFT, so Phi_w = G F (c(-.) Z S): grid pixel b sees
    the chirp image at pixel -b, whose generate_dde coordinates (wproj_utilities.cc:51-75) are
    X = (b - N/2 - 1/2) L / N, L = 2 sin(N cell / 2) (fov_cosines, wproj_utilities.cc:37-49);
  - a full chirp row has unit energy (Parseval);
  - for any G, a row is the reference's row-wise convolution (pinned by its known answers in
    test_oracle_wproj.py) of G's row with that chirp row;
  - the hard thresholds keep at least the requested energy fraction, within the bisection's 1 % band.
"""
import numpy as np
import pytest
from scipy.special import fresnel

import oracle

ARCSEC = np.pi / (180 * 3600)


def fresnel_transform_1d(a, w):
    """int exp(i pi w l^2 - 2 pi i a l) dl over the real line, from Fresnel integrals: completing the square
    it is exp(-i pi a^2 / w) int exp(i pi w t^2) dt, and with s = t sqrt(2|w|) the last integral is
    2 / sqrt(2|w|) (C(inf) + i sgn(w) S(inf)); C and S are taken at a large finite limit."""
    S, C = fresnel(1e9)
    return np.exp(-1j * np.pi * a * a / w) * 2.0 / np.sqrt(2.0 * abs(w)) * (C + 1j * np.sign(w) * S)


def screen_transform(a, b, w):
    return fresnel_transform_1d(a, w) * fresnel_transform_1d(b, w)


# ---- the approximate w-kernel --------------------------------------------------------------------------


def check_wkernel_is_minus_the_screen_transform(wk, cell, n, os_):
    du = 1.0 / (np.floor(os_ * n) * cell * ARCSEC)  # wavelengths per grid pixel
    rng = np.random.default_rng(1)
    for w in (50.0, -120.0, 200.0, -7.5):
        for _ in range(20):
            # inside the kernel's region |a du / w| < 1 on both axes
            a, b = rng.uniform(-0.95, 0.95, 2) * abs(w) / du
            got = wk(a, b, w, cell, cell, n, n, os_)
            want = -screen_transform(a * du, b * du, w)
            if abs(got - want) > 1e-7 * abs(want):
                return False
    return True


def test_wkernel_is_minus_the_fourier_transform_of_the_small_angle_screen():
    cell, n, os_ = 30.0, 64, 2.0
    assert check_wkernel_is_minus_the_screen_transform(oracle.wkernel, cell, n, os_)


@pytest.mark.parametrize("mistake", ["flipped exponent", "w^2 amplitude", "i w", "du x 2", "no -1"])
def test_the_screen_transform_check_catches_planted_mistakes(mistake):
    wk = oracle.wkernel
    mutants = {
        "flipped exponent": lambda a, b, w, *r: np.conj(wk(a, b, w, *r)) * (-1),  # exp(+i pi r^2/w)/(iw)
        "w^2 amplitude": lambda a, b, w, *r: wk(a, b, w, *r) / w,
        "i w": lambda a, b, w, *r: wk(a, b, w, *r) * -1,  # i/w in place of 1/(i w)
        "du x 2": lambda a, b, w, cx, cy, nx, ny, o: wk(a, b, w, cx / 2, cy / 2, nx, ny, o),
        "no -1": lambda a, b, w, *r: -wk(a, b, w, *r),
    }
    assert not check_wkernel_is_minus_the_screen_transform(mutants[mistake], 30.0, 64, 2.0)


def test_uv_cell_is_the_inverse_field_of_view_and_matches_the_unit_conversion():
    for cell, n, os_ in ((30.0, 64, 2.0), (1.7578125, 2048, 2.0), (60.0, 100, 1.5)):
        du = oracle.pixel_to_lambda(cell, n, os_)
        fov = np.floor(os_ * n) * cell * ARCSEC  # the grid's field of view in radians
        assert du == pytest.approx(1.0 / fov, rel=1e-14)
        # the operator's lambda -> pixel conversion (set_cell_size then uv_scale) uses the same cell
        u_l = np.array([123.4, -567.8, 1e4])
        u_px, v_px = oracle.uv_to_pixels(u_l, -u_l, 0, cell, cell, n, n, os_)
        np.testing.assert_allclose(u_l / u_px, du, rtol=1e-13)
        np.testing.assert_allclose(-u_l / v_px, du, rtol=1e-13)


def test_wkernel_cutoff_region():
    cell, n, os_ = 30.0, 64, 2.0
    du = oracle.pixel_to_lambda(cell, n, os_)
    for w in (80.0, -80.0):
        r = abs(w) / du  # pixels at which |u du / w| = 1
        inside = oracle.wkernel(0.999 * r, -0.999 * r, w, cell, cell, n, n, os_)
        assert inside == pytest.approx(-screen_transform(0.999 * r * du, -0.999 * r * du, w), rel=1e-7)
        for a, b in ((1.001 * r, 0.2 * r), (0.2 * r, -1.001 * r), (1.5 * r, 1.5 * r)):
            assert oracle.wkernel(a, b, w, cell, cell, n, n, os_) == 1.0
    # w = 0: every offset is outside the region (u du / 0), so the sample is 1
    assert oracle.wkernel(0.3, 0.7, 0.0, cell, cell, n, n, os_) == 1.0


def box_smear_windows(u, v, J, Jw, n=16, os_=2.0):
    """The integer shifts (su, sv) for which the box-w-kernel row of (u, v) equals the sum of the plain
    rows of (u + su, v + sv): found by least squares over candidate windows of Jw consecutive shifts."""
    one = np.ones(1, np.complex128)
    Gw = oracle.gridding_matrix_w([u], [v], [37.0], one, n, n, os_, oracle.KB, J, J, J, J, Jw, oracle.W_BOX)
    target = Gw.todense()[0]
    best = None
    for s0u in range(-Jw, 1):
        for s0v in range(-Jw, 1):
            acc = 0
            for su in range(s0u, s0u + Jw):
                for sv in range(s0v, s0v + Jw):
                    acc = acc + oracle.gridding_matrix([u + su], [v + sv], one, n, n, os_, oracle.KB, J, J, J,
                                                       J).todense()[0]
            err = np.linalg.norm(acc - target) / np.linalg.norm(target)
            if best is None or err < best[0]:
                best = (err, s0u, s0v)
    return best


@pytest.mark.parametrize("J,Jw", [(4, 3), (6, 5), (4, 4)])
def test_box_w_kernel_smears_the_row_over_shifts_centred_on_the_visibility(J, Jw):
    rng = np.random.default_rng(2)
    for _ in range(3):
        u, v = rng.uniform(-2, 2, 2)
        err, s0u, s0v = box_smear_windows(u, v, J, Jw)
        assert err < 1e-12
        cu, cv = s0u + (Jw - 1) / 2, s0v + (Jw - 1) / 2  # centre of each window of shifts
        if Jw % 2:
            assert cu == 0 and cv == 0
        else:
            assert abs(cu) == 0.5 and abs(cv) == 0.5


def test_operator_sees_w_minus_its_mean():
    n, M = 16, 200
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-8, 8, M), rng.uniform(-8, 8, M)
    w = rng.uniform(0, 200, M)
    x = rng.normal(size=n * n) + 1j * rng.normal(size=n * n)
    a = oracle.Operator(u, v, w, None, n, n, 2.0, oracle.KB, 4, 4, True, 4, 300.0, 300.0).forward(x)
    b = oracle.Operator(u, v, w + 17.25, None, n, n, 2.0, oracle.KB, 4, 4, True, 4, 300.0, 300.0).forward(x)
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)
    # and the rows do depend on w here: rows built from w itself (a planted "mean not subtracted") differ
    # from those of w - mean(w), so the invariance above is the subtraction's doing
    one = np.ones(M, np.complex128)
    raw = oracle.gridding_matrix_w(u, v, w, one, n, n, 2.0, oracle.KB, 4, 4, 4, 4, 4, oracle.W_APPROX, 300.0, 300.0)
    sub = oracle.gridding_matrix_w(u, v, w - w.mean(), one, n, n, 2.0, oracle.KB, 4, 4, 4, 4, 4, oracle.W_APPROX,
                                   300.0, 300.0)
    assert np.linalg.norm(raw.vals - sub.vals) > 1e-3 * np.linalg.norm(sub.vals)


# ---- chirp-based rows -------------------------------------------------------------------------------------


def dc_rows(M, N):
    return oracle.CSR(M, N * N, np.arange(M + 1, dtype=np.int64), np.zeros(M, np.int32), np.ones(M, np.complex128))


def screen(w, bp, bq, N, cell, shift=-0.5):
    L = 2 * np.sin(cell * N * ARCSEC / 2)
    X, Y = (bq - N / 2 + shift) * L / N, (bp - N / 2 + shift) * L / N
    return np.exp(-2j * np.pi * w * (np.sqrt(1 - X * X - Y * Y) - 1))


def point_source_errors(wproj, w_values, N=64, cell=600.0, shift=-0.5):
    """max over w and source pixels of |Phi_w delta_b - screen_b Phi delta_b| / |Phi delta_b| for the
    zero-frequency row (Phi delta_b = F delta_b at frequency 0 = 1 / N)."""
    M = len(w_values)
    GW = wproj(dc_rows(M, N), N, N, np.asarray(w_values, float), cell, cell, 1.0, 1.0)
    worst = 0.0
    for bp, bq in ((21, 39), (10, 50), (33, 31), (40, 12), (5, 60)):
        z = np.zeros(N * N, np.complex128)
        z[bp * N + bq] = 1.0
        yw = GW @ oracle.fft2d(N, N, z, -1)
        for m, w in enumerate(w_values):
            want = screen(w, bp, bq, N, cell, shift) / N
            worst = max(worst, abs(yw[m] - want) / abs(want))
    return worst


def oracle_wproj(G, x_size, y_size, w, cx, cy, ec, ew):
    return oracle.wprojection_matrix_dense(G, x_size, y_size, w, cx, cy, ec, ew)[0]


W_VALUES = [500.0, -300.0, 50.0, 1e3]


def test_chirp_rows_apply_the_exact_w_screen_to_a_point_source():
    assert point_source_errors(oracle_wproj, W_VALUES) < 1e-12


@pytest.mark.parametrize("mistake", ["chirp sign", "pixel centre", "flat field", "no sqrt(N)"])
def test_the_point_source_check_catches_planted_mistakes(mistake):
    if mistake == "chirp sign":  # exp(+2 pi i w ...): the oracle given -w
        err = point_source_errors(lambda G, x, y, w, *r: oracle_wproj(G, x, y, -w, *r), W_VALUES)
    elif mistake == "pixel centre":  # the chirp read at pixel b instead of -b (or an inverse FFT)
        err = point_source_errors(oracle_wproj, W_VALUES, shift=+0.5)
    elif mistake == "flat field":  # L = N cell instead of 2 sin(N cell / 2): the oracle given a smaller cell
        L = 2 * np.sin(600.0 * 64 * ARCSEC / 2)
        cell_flat = L / 64 / ARCSEC
        err = point_source_errors(lambda G, x, y, w, cx, cy, *r: oracle_wproj(G, x, y, w, cell_flat, cell_flat, *r),
                                  W_VALUES)
    else:  # the chirp row not scaled by sqrt(N)
        err = point_source_errors(lambda *a: oracle_wproj(*a) / 64.0, W_VALUES)
    assert err > 1e-6, err


def test_full_chirp_rows_have_unit_energy():
    N = 64
    GW = oracle_wproj(dc_rows(4, N), N, N, np.array([500.0, -300.0, 50.0, 0.0]), 600.0, 600.0, 1.0, 1.0)
    np.testing.assert_allclose(np.sum(np.abs(GW) ** 2, axis=1), 1.0, rtol=1e-12)


def test_rows_are_the_row_wise_convolution_of_g_with_the_chirp_row():
    n, N, cell = 32, 64, 600.0
    rng = np.random.default_rng(4)
    M = 4
    u, v = rng.uniform(-6, 6, M), rng.uniform(-6, 6, M)
    w = rng.uniform(-400, 400, M)
    G = oracle.gridding_matrix(u, v, np.ones(M, np.complex128), n, n, 2.0, oracle.KB, 4, 4, 4, 4)
    GW = oracle_wproj(G, N, N, w, cell, cell, 1.0, 1.0)
    chirp = oracle_wproj(dc_rows(M, N), N, N, w, cell, cell, 1.0, 1.0)
    for m in range(M):
        lo, hi = G.rowptr[m], G.rowptr[m + 1]
        want = oracle.row_wise_convolution(G.colidx[lo:hi], G.vals[lo:hi], chirp[m], N, N)
        np.testing.assert_allclose(GW[m], want, rtol=0, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("e_chirp,e_wproj", [(0.999, 1.0), (1.0, 0.99), (0.999, 0.999), (0.9, 0.95)])
def test_thresholds_keep_the_requested_energy(e_chirp, e_wproj):
    n, N, cell = 32, 64, 600.0
    rng = np.random.default_rng(5)
    M = 5
    u, v = rng.uniform(-6, 6, M), rng.uniform(-6, 6, M)
    w = rng.uniform(-400, 400, M)
    def check(cut, full, e):
        # at least the fraction e is kept, and the threshold is tight: the bisection either stopped inside
        # its 1 % band or (not converging in 200 steps) kept the largest threshold that still keeps e, so
        # dropping the smallest kept magnitude (every entry of that magnitude: the chirp's spectrum is
        # symmetric) would fall below e
        for c, f in zip(cut, full):
            tot = np.sum(np.abs(f) ** 2)
            kept = np.sum(np.abs(c) ** 2) / tot
            assert kept >= e * (1 - 1e-12)
            if e < 1:
                a = np.abs(c[c != 0])
                low = a <= a.min() * (1 + 1e-9)
                assert kept <= e * 1.01 + 1e-12 or kept - np.sum(a[low] ** 2) / tot < e
            nz = c != 0
            assert np.array_equal(c[nz], f[nz])  # thresholding only removes entries
            assert np.all(np.abs(f[~nz]) <= np.min(np.abs(c[nz])) + 1e-300)  # and removes the smallest

    # the chirp rows alone (unit-energy rows before the threshold)
    check(oracle_wproj(dc_rows(M, N), N, N, w, cell, cell, e_chirp, 1.0),
          oracle_wproj(dc_rows(M, N), N, N, w, cell, cell, 1.0, 1.0), e_chirp)
    # the final threshold of each convolved row against the same row before it
    G = oracle.gridding_matrix(u, v, np.ones(M, np.complex128), n, n, 2.0, oracle.KB, 4, 4, 4, 4)
    check(oracle_wproj(G, N, N, w, cell, cell, e_chirp, e_wproj), oracle_wproj(G, N, N, w, cell, cell, e_chirp, 1.0),
          e_wproj)
