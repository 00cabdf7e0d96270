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


SARA3 = [(0, 1), (1, 2), (4, 2)]  # Dirac, DB1, DB4


def _ops(n, ents):
    return (lambda v: oracle.sara_analysis(v, n, n, ents)), (lambda c: oracle.sara_synthesis(c, n, n, ents))


def test_iterative_proximal_without_constraints_reaches_the_tight_frame_closed_form():
    # one orthonormal basis, no constraints: the minimiser is Psi soft(Psi^dagger x0, gamma)
    n, ents = 16, [(4, 2)]
    psit, psi = _ops(n, ents)
    rng = np.random.default_rng(11)
    x0 = rng.normal(size=n * n) + 1j * rng.normal(size=n * n)
    g = 0.4
    x, k = oracle.l1_prox_constrained(x0, g, psit, psi, 1.0, 500, 0.0, positivity=False, real=False)
    assert np.allclose(x, psi(oracle.soft_threshold(psit(x0), g)), atol=1e-10)


def _objective(x, x0, g, psit):
    return 0.5 * np.vdot(x - x0, x - x0).real + g * np.abs(psit(x)).sum()


def test_iterative_proximal_matches_an_independent_admm_solver_on_sara_with_constraints():
    # the objective is strongly convex, so its minimiser is unique: a plain ADMM on the split
    # min 1/2||x - x0||^2 + i_C(x) + gamma ||w||_1 s.t. w = Psi^dagger x (Psi Psi^dagger = I for SARA)
    # must reach the same point
    n = 8
    psit, psi = _ops(n, SARA3)
    rng = np.random.default_rng(12)
    x0 = rng.normal(size=n * n) + 0.3 * rng.normal(size=n * n) * 1j
    g = 0.15
    x, k = oracle.l1_prox_constrained(x0, g, psit, psi, 1.0, 20000, 0.0)
    rho = 1.0
    w = psit(oracle.project_c(x0))
    lam = np.zeros_like(w)
    xa = oracle.project_c(x0)
    for _ in range(20000):
        # x-step: argmin 1/2||x - x0||^2 + i_C + rho/2 ||Psi^dagger x - w + lam||^2; with Psi Psi^dagger = I
        # this is P_C((x0 + rho Psi(w - lam)) / (1 + rho))
        xa = oracle.project_c((x0 + rho * psi(w - lam)) / (1 + rho))
        cw = psit(xa) + lam
        w = oracle.soft_threshold(cw, g / rho)
        lam = lam + psit(xa) - w
    assert np.allclose(xa.imag, 0) and np.all(xa.real >= 0)
    assert np.allclose(x, xa, atol=1e-6)
    # and it is no worse than nearby feasible points
    f0 = _objective(x, x0, g, psit)
    for _ in range(20):
        p = oracle.project_c(x + 1e-3 * rng.normal(size=n * n))
        assert _objective(p, x0, g, psit) >= f0 - 1e-12


def test_iterative_proximal_stops_on_the_objective_change():
    n = 16
    psit, psi = _ops(n, SARA3)
    rng = np.random.default_rng(13)
    x0 = rng.normal(size=n * n)
    _, k_loose = oracle.l1_prox_constrained(x0, 0.2, psit, psi, 1.0, 50, 1e-2)
    _, k_tight = oracle.l1_prox_constrained(x0, 0.2, psit, psi, 1.0, 50, 1e-8)
    _, k_cap = oracle.l1_prox_constrained(x0, 0.2, psit, psi, 1.0, 3, 0.0)
    assert 1 <= k_loose < k_tight and k_cap == 3


def test_sara_padmm_reconstruction_improves_on_the_dirty_image():
    op, x0, y, sigma = small(5, 32, 2000)
    eps = utilities.calculate_l2_radius(y.size, sigma)
    x, st, g, it = oracle.padmm(op, [(0, 1), (2, 2), (4, 2)], y, eps, 60, tight_frame=False)
    dirty = op.adjoint(y)
    dirty = dirty * np.vdot(dirty, x0).real / np.vdot(dirty, dirty).real
    assert oracle.snr_db(x0, x) > oracle.snr_db(x0, dirty)
    assert np.allclose(x.imag, 0) and np.all(x.real >= 0)


def test_forward_backward_without_regularisation_is_landweber_on_the_dense_matrix():
    # mu -> 0 with the closed-form proximal: x_(k+1) = x_k - step (A^H A x_k - A^H y), checked against the
    # dense matrix of the same operator
    n, M = 8, 40
    rng = np.random.default_rng(21)
    u, v = rng.normal(0, 0.3, M) * n, rng.normal(0, 0.3, M) * n
    op = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, 4, 4, power_iters=50, power_tol=1e-10)
    A = np.stack([op.forward(np.eye(n * n)[k].astype(np.complex128)) for k in range(n * n)], axis=1)
    y = rng.normal(size=M) + 1j * rng.normal(size=M)
    sigma, beta = 0.7, 0.3
    x, st = oracle.forward_backward(op, [(4, 1)], y, 15, sigma, mu=1e-300, beta=beta)
    xd = np.zeros(n * n, np.complex128)
    for _ in range(15):
        xd = xd - beta / sigma ** 2 * (A.conj().T @ (A @ xd) - A.conj().T @ y)
    assert np.allclose(x, xd, atol=1e-12)


def test_forward_backward_decreases_its_objective_and_reaches_a_fixed_point():
    op, x0, y, sigma = small(7, 16, 400)
    ents = [(4, 2)]
    psit = lambda v: oracle.sara_analysis(v, 16, 16, ents)  # noqa: E731
    mu, beta = 2.0, sigma ** 2  # step beta / sigma^2 = 1 <= 1 / ||Phi||^2
    F = lambda x: 0.5 * np.linalg.norm(y - op.forward(x)) ** 2 / sigma ** 2 + mu * np.abs(psit(x)).sum()  # noqa: E731
    vals = [F(oracle.forward_backward(op, ents, y, k, sigma, mu, beta)[0]) for k in (1, 2, 4, 8, 16)]
    assert all(b <= a + 1e-9 * abs(a) for a, b in zip(vals, vals[1:]))
    # the FB map is averaged for step <= 1 / L, so ||x_(k+1) - x_k|| never grows and tends to 0
    x, st = oracle.forward_backward(op, ents, y, 400, sigma, mu, beta)
    d = np.sqrt(st[:, 0])
    assert np.all(d[2:] <= d[1:-1] * (1 + 1e-9))
    assert d[-1] < 1e-2 * d[1]
