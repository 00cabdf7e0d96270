"""CPU pins of the oracle PADMM: invariants the iteration must satisfy whatever SOPT's internals
(which are not in the reference tree, so parity with SOPT's iterates is unpinned)."""
import numpy as np

import oracle
from purify_amd import utilities


def small(seed=1, n=16, M=300):
    rng = np.random.default_rng(seed)
    u, v = rng.normal(0, 0.3, M) * n, rng.normal(0, 0.3, M) * n
    op = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, 4, 4, power_iters=50, power_tol=1e-8)
    x0 = utilities.synth_image(seed + 1, n, n, 5)
    y0 = op.forward(x0)
    sigma = utilities.SNR_to_standard_deviation(y0, 30.0)
    return op, x0, utilities.add_noise(y0, sigma, seed + 2), sigma


def test_unconstrained_and_unregularised_iteration_is_stationary():
    # a ball containing every residual makes z = -lambda - r, so x never moves from Phi^dagger y / nu and
    # lambda stays 0; with gamma -> 0 the l1 proximal is the identity
    op, x0, y, _ = small()
    x, st, g, it = oracle.padmm(op, [(4, 1)], y, 1e12, 5, gamma=1e-300)
    assert np.allclose(x, op.adjoint(y), atol=1e-12)
    assert np.all(st[:, 0] < 1e-24)


def test_tight_frame_proximal_is_the_synthesis_of_the_thresholded_analysis():
    n = 16
    rng = np.random.default_rng(3)
    xt = rng.normal(size=n * n) + 1j * rng.normal(size=n * n)
    ents = [(4, 2)]
    c = oracle.sara_analysis(xt, n, n, ents)
    a = xt + oracle.sara_synthesis(oracle.soft_threshold(c, 0.3) - c, n, n, ents)
    b = oracle.sara_synthesis(oracle.soft_threshold(c, 0.3), n, n, ents)
    assert np.allclose(a, b, atol=1e-12)
    # soft threshold: magnitudes shrink by t, phases stay, |c| <= t vanish
    s = oracle.soft_threshold(np.array([3 + 4j, 0.1j, -2.0]), 1.0)
    assert np.allclose(s, [(3 + 4j) * 0.8, 0, -1.0])


def test_reconstruction_improves_on_the_dirty_image_and_converges():
    op, x0, y, sigma = small(5, 32, 2000)
    eps = utilities.calculate_l2_radius(y.size, sigma)
    x, st, g, it = oracle.padmm(op, [(4, 2)], y, eps, 300, stop=True)
    dirty = op.adjoint(y)
    dirty = dirty * np.vdot(dirty, x0).real / np.vdot(dirty, dirty).real  # best scaling of the dirty image
    assert oracle.snr_db(x0, x) > oracle.snr_db(x0, dirty)
    assert it < 300 and np.sqrt(st[-1, 2]) < eps
