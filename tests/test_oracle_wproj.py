"""Pins the oracle's w-projection pieces against the reference's tests.

Reference tests restated:
  cpp/tests/convolution.cc:35-148   linear_convol_2d known answers (box/delta)
  cpp/tests/wproj_utils.cc:43-83    a Jw=1 box w-kernel reproduces the no-w G
  cpp/tests/wproj_utils.cc:127-136  chirp at w = 0 has |c| = 1/N, 1e-12
  cpp/tests/wproj_utils.cc:138-173  threshold keeps the requested energy within 10 %
  cpp/tests/wproj_utils.cc:187-216  wprojection_matrix with w = 0 is the identity, 1e-12
  cpp/tests/wproj_utils.cc:217-289  row-wise convolution known answers, 1e-7
"""
import numpy as np
import pytest

import oracle
from golden import fixtures


def isapprox(a, b, prec):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


def test_convol_box_and_delta():
    Ju, Jv, Jw = 3, 4, 5
    fu = np.ones(Ju, np.complex128)
    fv = np.ones(Jv, np.complex128)
    g = np.zeros((Jw, Jw), np.complex128)
    g[2, 2] = 1.0  # floor(j - (Jw-1)/2) == 0 only at j = 2
    out = oracle.linear_convol_2d(fu, fv, g)
    assert out.shape == (Jv + Jw - 1, Ju + Jw - 1)
    assert isapprox(out[2:2 + Jv, 2:2 + Ju], np.ones((Jv, Ju)), 1e-12)


def test_convol_delta_and_delta():
    Ju, Jv, Jw = 3, 4, 5
    fu = np.zeros(Ju, np.complex128)
    fu[1] = 1.0
    fv = np.zeros(Jv, np.complex128)
    fv[2] = 1.0  # floor(jv - (Jv-1)/2) == 0 at jv = 2 for Jv = 4
    g = np.zeros((Jw, Jw), np.complex128)
    g[2, 2] = 1.0
    out = oracle.linear_convol_2d(fu, fv, g)
    assert out[4, 3] == 1.0


def test_convol_box_and_delta_size_1():
    Ju, Jv = 3, 4
    fu = np.arange(Ju, dtype=np.complex128)
    fv = np.arange(Jv, dtype=np.complex128)
    g = np.ones((1, 1), np.complex128)
    out = oracle.linear_convol_2d(fu, fv, g)
    for i in range(out.shape[1]):
        for j in range(out.shape[0]):
            assert abs(i * j - out[j, i]) < 1e-12


def test_convol_2d_against_numpy_convolve():
    # linear_convol_1d(f, pad(g)) is the full convolution of f with reversed g
    rng = np.random.default_rng(8)
    f = rng.normal(size=4) + 1j * rng.normal(size=4)
    g = rng.normal(size=7) + 1j * rng.normal(size=7)
    padded = np.concatenate([np.zeros(3), g, np.zeros(3)])
    np.testing.assert_allclose(oracle.linear_convol_1d(f, padded), np.convolve(f, g[::-1]), rtol=1e-13)


def test_box_wkernel_size_one_reproduces_plain_gridding():
    rng = np.random.default_rng(9)
    Ju = Jv = 5
    n, M = 10, 5
    u = rng.uniform(-1, 1, M)
    v = rng.uniform(-1, 1, M)
    w = rng.uniform(-1, 1, M) * 100
    wts = np.ones(M, np.complex128)
    Gw = oracle.gridding_matrix_w(u, v, w, wts, n, n, 2.0, oracle.KB, Ju, Jv, Ju, Jv, 1, oracle.W_BOX)
    G = oracle.gridding_matrix(u, v, wts, n, n, 2.0, oracle.KB, Ju, Jv, Ju, Jv)
    assert Gw.nnz == G.nnz
    assert isapprox(Gw.todense(), G.todense(), 1e-12)


def test_approx_wkernel_runs_and_is_finite():
    rng = np.random.default_rng(10)
    n, M, J, Jw = 10, 5, 5, 10
    u, v = rng.uniform(-1, 1, M), rng.uniform(-1, 1, M)
    w = rng.uniform(-1, 1, M) * 100
    Gw = oracle.gridding_matrix_w(u, v, w, np.ones(M, np.complex128), n, n, 2.0, oracle.KB, J, J, J, J, Jw,
                                  oracle.W_APPROX, 30.0, 30.0)
    assert Gw.nnz == M * (J + Jw - 1) ** 2
    assert np.all(np.isfinite(Gw.vals))


def test_chirp_w_zero_has_modulus_one_over_n():
    c = oracle.generate_chirp(0.0, 1.0, 1.0, 128, 128)
    assert np.linalg.norm(np.abs(c) - 1.0 / (128 * 128)) < 1e-12


@pytest.mark.parametrize("energy", [1.0, 0.4])
def test_threshold_energy(energy):
    rng = np.random.default_rng(11)
    row = rng.uniform(-1, 1, 100_000)
    thr = oracle.sparsify_row_thres(np.abs(row), energy)
    total = np.sum(row ** 2)
    kept = np.sum(row[np.abs(row) > thr] ** 2) / total
    assert abs(energy - kept) / energy < 1e-1


def test_row_wise_convolution_known_answers():
    ka = fixtures.convolution_known_answers()
    for key in ("even_4x4", "odd_5x5"):
        d = ka[key]
        N = d["N"]
        G = np.array(d["G"], np.complex128)
        Cc = np.array(d["C"], np.complex128)
        K = np.array(d["K"], np.complex128)
        idx = np.nonzero(np.abs(G) > 1e-10)[0].astype(np.int32)
        out = oracle.row_wise_convolution(idx, G[idx], Cc, N, N)
        assert isapprox(out, K, 1e-7)


def test_row_wise_convolution_diracs():
    N = 4
    G = np.zeros(N * N, np.complex128)
    G[0] = 1
    Cc = np.zeros(N * N, np.complex128)
    Cc[5] = 1
    out = oracle.row_wise_convolution(np.array([0], np.int32), G[:1], Cc, N, N)
    assert isapprox(out, Cc, 1e-7)
    N = 5
    rng = np.random.default_rng(12)
    G = rng.uniform(-1, 1, N * N) + 1j * rng.uniform(-1, 1, N * N)
    Cc = np.zeros(N * N, np.complex128)
    Cc[0] = 1
    out = oracle.row_wise_convolution(np.arange(N * N, dtype=np.int32), G, Cc, N, N)
    assert isapprox(out, G, 1e-7)


def test_wprojection_matrix_identity_at_w_zero():
    rows, N = 12, 32
    rowptr = np.arange(rows + 1, dtype=np.int64)
    colidx = np.arange(rows, dtype=np.int32)
    I = oracle.CSR(rows, N * N, rowptr, colidx, np.ones(rows, np.complex128))
    rng = np.random.default_rng(13)
    dense0, mask0 = oracle.wprojection_matrix_dense(I, N, N, np.zeros(rows), 60.0, 60.0, 1.0, 1.0)
    assert mask0.sum() == rows
    for k in range(rows):
        nz = np.nonzero(mask0[k])[0]
        assert list(nz) == [k]
        assert abs(dense0[k, k] - 1.0) < 1e-12
    dense, mask = oracle.wprojection_matrix_dense(I, N, N, rng.uniform(-1, 1, rows), 60.0, 60.0, 1.0, 1.0)
    assert mask.sum() >= rows
