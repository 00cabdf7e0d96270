"""CPU pins of the wavelet dictionary (sopt::wavelets as wavelet_operator_factory.h:20-45 uses it).

SOPT is not in the reference tree: parity against SOPT's own coefficients and transform is unpinned.
What is pinned: the library's and the oracle's Daubechies filters against their defining identities
(sum sqrt(2), orthonormal shifts, vanishing moments), the Haar and DB2 closed forms and the published
DB4 filter; the oracle transform against orthonormality, perfect reconstruction and a Haar known
answer."""
import numpy as np
import pytest

import oracle
from purify_amd import wavelets


@pytest.mark.parametrize("N", range(1, 11))
def test_daubechies_filters_satisfy_their_identities(N):
    for h in (wavelets.daubechies_filter(N), oracle.daubechies(N)):
        assert len(h) == 2 * N
        assert h.sum() == pytest.approx(np.sqrt(2.0), abs=1e-13)
        for m in range(N):  # sum_k h[k] h[k + 2m] = delta_m
            assert np.dot(h[: len(h) - 2 * m], h[2 * m:]) == pytest.approx(1.0 if m == 0 else 0.0, abs=1e-12)
        k = np.arange(2 * N)
        for p in range(N):  # vanishing moments of the high-pass filter
            assert np.sum((-1.0) ** k * k ** p * h) == pytest.approx(0.0, abs=1e-9 * (2 * N) ** p)
    assert np.allclose(wavelets.daubechies_filter(N), oracle.daubechies(N), rtol=0, atol=1e-13)


def test_haar_db2_and_published_db4():
    assert np.allclose(wavelets.daubechies_filter(1), [2 ** -0.5, 2 ** -0.5], atol=1e-16)
    r3 = np.sqrt(3.0)
    db2 = np.array([1 + r3, 3 + r3, 3 - r3, 1 - r3]) / (4 * np.sqrt(2.0))
    assert np.allclose(wavelets.daubechies_filter(2), db2, atol=1e-15)
    db4 = [0.2303778133088964, 0.7148465705529154, 0.6308807679298587, -0.0279837694168599,
           -0.1870348117190931, 0.0308413818355607, 0.0328830116668852, -0.0105974017850690]
    assert np.allclose(wavelets.daubechies_filter(4), db4, atol=1e-13)


def test_oracle_transform_is_orthonormal_and_inverts():
    ny, nx = 32, 16
    rng = np.random.default_rng(5)
    x = rng.normal(size=ny * nx) + 1j * rng.normal(size=ny * nx)
    for N, lev in ((1, 3), (4, 2), (8, 1), (3, 4)):
        h = oracle.daubechies(N)
        c = oracle.dwt2(x, ny, nx, h, lev)
        assert np.linalg.norm(c) == pytest.approx(np.linalg.norm(x), rel=1e-13)
        assert np.allclose(oracle.idwt2(c, ny, nx, h, lev), x, atol=1e-12)
    entries = [(0, 3)] + [(k, 3) for k in range(1, 9)]  # the reference test's SARA (Dirac, DB1..DB8)
    c = oracle.sara_analysis(x, ny, nx, entries)
    assert c.size == 9 * ny * nx
    assert np.allclose(oracle.sara_synthesis(c, ny, nx, entries), x, atol=1e-12)  # a tight frame


def test_haar_one_level_known_answer():
    # 2 x 2 image: one Haar level gives (a + b + c + d) / 2 and the three differences / 2
    a, b, c, d = 1.0, 2.0, 5.0, 7.0
    out = oracle.dwt2(np.array([a, b, c, d]), 2, 2, oracle.daubechies(1), 1).reshape(2, 2)
    assert np.allclose(out, [[(a + b + c + d) / 2, (a - b + c - d) / 2], [(a + b - c - d) / 2, (a - b - c + d) / 2]])


def test_factory_errors():
    with pytest.raises(ValueError):
        wavelets.parse_wavelet("DB11")
    with pytest.raises(ValueError):
        wavelets.parse_wavelet("sym4")
    assert wavelets.parse_wavelet("Dirac") == 0 and wavelets.parse_wavelet("DB8") == 8
    with pytest.raises(RuntimeError, match="MPI"):
        wavelets.wavelet_operator_factory("mpi_sara", [("DB4", 3)], 16, 16)


@pytest.mark.parametrize("order,levels,ny,nx", [(1, 1, 8, 8), (4, 3, 64, 32), (8, 2, 16, 64), (10, 4, 128, 128),
                                                (2, 5, 64, 64)])
def test_oracle_c_transform_equals_the_numpy_restatement(order, levels, ny, nx):
    # the C lines (orc_dwt2 / orc_idwt2) and the numpy lines compute the same periodic transform
    rng = np.random.default_rng(order * 100 + levels)
    x = rng.normal(size=ny * nx) + 1j * rng.normal(size=ny * nx)
    h = oracle.daubechies(order)
    a, b = oracle.dwt2(x, ny, nx, h, levels), oracle.dwt2_numpy(x, ny, nx, h, levels)
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    c, d = oracle.idwt2(x, ny, nx, h, levels), oracle.idwt2_numpy(x, ny, nx, h, levels)
    assert np.abs(c - d).max() < 1e-13 * np.abs(d).max()
    assert np.abs(oracle.idwt2(a, ny, nx, h, levels) - x).max() < 1e-12
