"""Pins the oracle's scalar kernels and correction image against the reference fixtures.

Reference tests restated: cpp/tests/kernels.cc:35-73 (abs 1e-13) and cpp/tests/operators.cc:93-101
(S, Eigen isApprox 1e-6).  I0 is additionally pinned against scipy's Bessel routine.
"""
import numpy as np
import pytest
import scipy.special

import oracle
from golden import fixtures


def isapprox(a, b, prec):
    """Eigen's isApprox: ||a - b|| <= prec * min(||a||, ||b||)."""
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


def test_bessel_i0_matches_scipy():
    x = np.linspace(0, 30, 301)
    ours = oracle.bessel_i0(x)
    np.testing.assert_allclose(ours, scipy.special.i0(x), rtol=2e-15, atol=0)


@pytest.mark.parametrize("name,fn", [("kb", oracle.kaiser_bessel), ("gauss", oracle.gaussian),
                                     ("pswf", oracle.pswf)])
def test_kernel_fixture(name, fn):
    x = fixtures.read_reals("kernels", "kernel_x_data")
    expected = fixtures.read_reals("kernels", f"kernel_{name}_data")
    assert len(x) == len(expected) == 601
    got = fn(x, 6)
    assert np.max(np.abs(got - expected)) < 1e-13


@pytest.mark.parametrize("name,fn", [("kb", oracle.ft_kaiser_bessel), ("gauss", oracle.ft_gaussian),
                                     ("pswf", oracle.ft_pswf)])
def test_correction_fixture(name, fn):
    x = fixtures.read_reals("kernels", "correction_x_data")
    expected = fixtures.read_reals("kernels", f"correction_{name}_data")
    assert len(x) == len(expected) == 201
    got = fn(x, 6)
    assert np.max(np.abs(got - expected)) < 1e-13


def test_kb_endpoints_closed_form():
    # KB(+-J/2) = 1/I0(alpha), KB(0) = 1 (kernels.cc:15-22 closed form at the support edge)
    for J in (4, 6, 8):
        a = 2.34 * J
        assert oracle.kaiser_bessel([0.0], J)[0] == pytest.approx(1.0, rel=1e-15)
        assert oracle.kaiser_bessel([J / 2], J)[0] == pytest.approx(1.0 / scipy.special.i0(a), rel=1e-14)


def test_correction_image_fixture():
    expected = fixtures.read_reals("operators", "expected_S_data")
    S = oracle.correction2d(2.0, 16, 16, oracle.KB, 4, 4)
    assert len(expected) == 256
    assert isapprox(S, expected, 1e-6)


def test_correction_is_separable_outer_product():
    S = oracle.correction2d(2.0, 12, 20, oracle.KB, 4, 6).reshape(12, 20)
    # rank one: every row is a multiple of the first
    u, s, vt = np.linalg.svd(S)
    assert s[1] / s[0] < 1e-14


def test_pswf_requires_support_six():
    with pytest.raises(RuntimeError):
        oracle.correction2d(2.0, 8, 8, oracle.PSWF, 4, 4)
