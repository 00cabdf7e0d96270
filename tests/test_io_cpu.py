"""FITS image and .vis I/O (pfitsio.cc:34-211, utilities.cc:140-202) on the reference's own fixtures:
M31.fits and the PADMM factory test's input_data.vis / solution.fits (cpp/tests/algo_factory.cc:20-67)."""
import os

import numpy as np
import pytest

from purify_amd import pfitsio, utilities

D = os.path.join(os.path.dirname(__file__), "golden", "reference_data")
PADMM = os.path.join(D, "padmm")


def test_read_visibility_follows_the_reference_conventions():
    uv = utilities.read_visibility(os.path.join(PADMM, "input_data.vis"), False)
    assert uv.size() == 13107  # algo_factory.cc:35
    first = [float(t) for t in open(os.path.join(PADMM, "input_data.vis")).readline().split()]
    assert uv.u[0] == first[0] and uv.v[0] == -first[1]  # v reflected (utilities.cc:175)
    assert uv.vis[0] == complex(first[2], first[3]) and uv.weights[0] == 1.0 / first[4]
    assert np.all(uv.w == 0)


def test_write_visibility_round_trips(tmp_path):
    rng = np.random.default_rng(4)
    M = 50
    uv = utilities.vis_params(u=rng.normal(size=M), v=rng.normal(size=M), w=rng.uniform(0, 100, M),
                              vis=rng.normal(size=M) + 1j * rng.normal(size=M), weights=1.0 / rng.uniform(0.5, 2, M))
    for w_term in (False, True):
        p = tmp_path / f"t{int(w_term)}.vis"
        utilities.write_visibility(uv, str(p), w_term)
        assert len(open(p).readline().split()) == (6 if w_term else 5)
        back = utilities.read_visibility(str(p), w_term)
        assert np.allclose(back.u, uv.u, rtol=1e-12) and np.allclose(back.v, uv.v, rtol=1e-12)
        assert np.allclose(back.vis, uv.vis, rtol=1e-12) and np.allclose(back.weights, uv.weights, rtol=1e-12)
        assert np.allclose(back.w, uv.w if w_term else 0, rtol=1e-12)


def test_read2d_of_the_reference_images():
    m31 = pfitsio.read2d(os.path.join(D, "images", "M31.fits"))
    assert m31.shape == (256, 256) and m31.dtype == np.complex128 and np.all(m31.imag == 0)
    sol = pfitsio.read2d(os.path.join(PADMM, "solution.fits"))
    h = pfitsio.read_header(os.path.join(PADMM, "solution.fits"))
    assert sol.shape == (h["NAXIS2"], h["NAXIS1"]) == (256, 256)
    assert h["CRPIX1"] == 129 and h["CDELT1"] < 0  # the reference's WCS keywords (pfitsio.cc:119-133)
    assert np.all(np.isfinite(sol.real)) and sol.real.max() > 0


def test_write2d_round_trips_and_keeps_the_flat_order(tmp_path):
    rng = np.random.default_rng(5)
    a = rng.normal(size=(24, 40))  # NAXIS2 = 24, NAXIS1 = 40
    p = str(tmp_path / "a.fits")
    hp = pfitsio.header_params(fits_name=p, cell_x=2.0, cell_y=3.0, niters=17, hasconverged=True, epsilon=0.5)
    pfitsio.write2d(a, hp)
    assert os.path.getsize(p) % 2880 == 0
    b = pfitsio.read2d(p)
    assert b.shape == a.shape and np.array_equal(b.real, a) and np.array_equal(b.ravel().real, a.ravel())
    h = pfitsio.read_header(p)
    assert (h["NAXIS"], h["NAXIS1"], h["NAXIS2"], h["NAXIS3"], h["NAXIS4"]) == (4, 40, 24, 1, 1)
    assert h["CRPIX1"] == 21 and h["CRPIX2"] == 13
    assert h["CDELT1"] == pytest.approx(-2.0 / 3600) and h["CDELT2"] == pytest.approx(3.0 / 3600)
    assert h["BUNIT"] == "Jy/Beam" and h["CTYPE1"] == "RA--SIN"
    text = open(p, "rb").read(2880 * 2).decode("ascii", "replace")
    assert "PURIFY-NITERS: 17" in text and "PURIFY-CONVERGED: True" in text
    # cube: planes in order; overwrite refused when asked
    pfitsio.write3d([a, 2 * a], p, "K")
    planes = pfitsio.read3d(p)
    assert len(planes) == 2 and np.array_equal(planes[1].real, 2 * a)
    with pytest.raises(RuntimeError):
        pfitsio.write2d(a, p, overwrite=False)
