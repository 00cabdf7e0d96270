"""Pins of the oracle's compact operators (compact_operators.h): the explicit, pruned G^H G against the
operator's own G^H(G .), the PSF convolution against a brute-force circular convolution, and the
power iteration against closed-form eigenvalues."""
import numpy as np
import pytest

import oracle


def uv(seed, M, n, sigma=0.35):
    rng = np.random.default_rng(seed)
    return rng.normal(0, sigma, M) * n, rng.normal(0, sigma, M) * n


def cnormal(seed, n):
    rng = np.random.default_rng(seed)
    return rng.normal(size=n) + 1j * rng.normal(size=n)


@pytest.mark.parametrize("J", [4, 6])
def test_unpruned_gtg_equals_the_operator_composition(J):
    n, M = 16, 300
    u, v = uv(1 + J, M, n)
    w = cnormal(2, M)
    op = oracle.Operator(u, v, None, w, n, n, 2.0, oracle.KB, J, J)
    op.set_norm(1.7)
    x = cnormal(3, n * n)
    a = oracle.grid_degrid(u, v, w, n, n, 2.0, x, oracle.KB, J, J, norm=1.7, prune=0)
    b = op.adjoint(op.forward(x))
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-13


def test_pruning_drops_only_small_entries_and_changes_little():
    n, M = 16, 400
    u, v = uv(11, M, n)
    G = oracle.gridding_matrix(u, v, np.ones(M, np.complex128), n, n, 2.0, oracle.KB, 4, 4, 4, 4)
    full = oracle.gtg_matrix(G, prune=0)
    pruned = oracle.gtg_matrix(G, prune=1e-7)
    assert pruned.nnz < full.nnz
    d = (full - pruned).tocoo()
    assert np.abs(d.data).max() <= 1e-7
    # Hermitian (and so is the operator built on it)
    assert abs(full - full.conj().T).max() < 1e-15
    x = cnormal(12, n * n)
    a = oracle.grid_degrid(u, v, None, n, n, 2.0, x, prune=1e-7)
    b = oracle.grid_degrid(u, v, None, n, n, 2.0, x, prune=0)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-6


def test_psf_convolution_is_a_circular_convolution_by_brute_force():
    n, M = 8, 60
    u, v = uv(21, M, n)
    op = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, 4, 4)
    x = cnormal(22, n * n)
    got = oracle.psf_convolve(op, x)
    delta = np.zeros(n * n, np.complex128)
    delta[(n // 2 - 1) * n + n // 2 - 1] = 1
    psf = op.adjoint(op.forward(delta)).reshape(n, n)
    X = x.reshape(n, n)
    ref = np.zeros((n, n), np.complex128)
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    ref[i, j] += psf[k, l] * X[(i - k) % n, (j - l) % n]
    ref /= n  # unitary FFTs: FFT / sqrt(N) both ways, sqrt(N) = n
    assert np.allclose(got.reshape(n, n), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_power_iteration_closed_form():
    d = np.array([3.0, 2.0, 0.5, 0.1])
    lam = oracle.power_method_fn(lambda x: d * x, np.ones(4, np.complex128), 500, 1e-8)
    assert lam == pytest.approx(np.sqrt(3.0), rel=1e-6)
    assert oracle.power_method_fn(lambda x: x, np.ones(4), 0, 1e-4) == 1.0
