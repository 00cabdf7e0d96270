"""Operator state changes and the boundary's argument checks on the GPU:

- a solver whose iteration was captured into a graph captures again after pf_op_set_norm,
  pf_op_set_tuning or pf_op_autotune, so its replays equal eager iterations of the changed operator;
- pf_op_autotune only picks gridding kernels with a fixed summation order: the complex128 operator at the
  2048^2 / 4e6 shape autotunes to one and gives bit-identical adjoints;
- the tight-frame l1 proximal takes its own frame bound l1_nu, separate from the ADMM step nu (SOPT's
  L1TightFrame nu), checked against the oracle with nu = 2, l1_nu = 1;
- device tensors of the wrong size, dtype or layout are refused in Python before the C library is called.
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import make_uv, rel
from purify_amd import algorithms, measurementoperator, utilities, wavelets

pytestmark = pytest.mark.gpu

DB4 = [("DB4", 3)]


def problem(seed, n, M, snr=30.0, J=4):
    u, v = make_uv(seed, M, n)
    x0 = utilities.synth_image(seed + 1, n, n, 20)
    orc = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, J, J, power_iters=100, power_tol=1e-8)
    y0 = orc.forward(x0)
    sigma = utilities.SNR_to_standard_deviation(y0, snr)
    y = utilities.add_noise(y0, sigma, seed + 2)
    eps = utilities.calculate_l2_radius(M, sigma)
    return u, v, x0, y, eps, orc


def solver(u, v, y, n, eps, precision, norm, iters, **kw):
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                     precision=precision)
    lt.op.set_norm(norm)
    W = wavelets.WaveletOperator(DB4, n, n, precision)
    return algorithms.ImagingProximalADMM(lt.op, W, y, eps, itermax=iters, **kw), lt, W


@pytest.mark.parametrize("change", ["norm", "tuning", "autotune"])
def test_graph_is_captured_again_after_the_operator_changes(change):
    n, M = 64, 4000
    u, v, x0, y, eps, orc = problem(1101, n, M)
    a, lta, _wa = solver(u, v, y, n, eps, "c64", orc.norm, 40)
    b, ltb, _wb = solver(u, v, y, n, eps, "c64", orc.norm, 40)
    a.iterate(3, use_graph=False)
    b.iterate(3, use_graph=True)  # captured with the first norm and tuning
    assert np.array_equal(a.image(), b.image())
    if change == "norm":
        for lt in (lta, ltb):
            lt.op.set_norm(orc.norm * 1.7)
    elif change == "tuning":
        for lt in (lta, ltb):
            lt.op.set_tuning("grid_kernel", 0)
            lt.op.set_tuning("fft_columns", 1)
    else:
        ltb.op.set_tuning("grid_kernel", 0)
        ltb.op.autotune(2)  # re-plans the column FFT: its twiddle table is allocated anew
        for key in ("grid_kernel", "fft_columns"):
            lta.op.set_tuning(key, ltb.op.get_tuning(key))
    for _ in range(3):
        a.iterate(3, use_graph=False)
        b.iterate(3, use_graph=True)
        torch.cuda.synchronize()
        assert np.array_equal(a.image(), b.image())
        assert np.array_equal(a.residual(), b.residual())
    assert np.array_equal(a.stats(), b.stats())
    if change == "norm":
        # and the changed norm is the one the iterations used: not those of the first operator
        c, ltc, _wc = solver(u, v, y, n, eps, "c64", orc.norm, 40)
        c.iterate(12, use_graph=False)
        assert not np.array_equal(c.image(), b.image())


def test_reset_after_a_norm_change_recomputes_gamma_for_the_graph():
    n, M = 64, 4000
    u, v, x0, y, eps, orc = problem(1111, n, M)
    a, lta, _wa = solver(u, v, y, n, eps, "c128", orc.norm, 40)
    b, ltb, _wb = solver(u, v, y, n, eps, "c128", orc.norm, 40)
    b.iterate(2, use_graph=True)
    for s, lt in ((a, lta), (b, ltb)):
        lt.op.set_norm(orc.norm * 0.5)
        s.reset()
    ra = a.iterate(6, use_graph=False)
    rb = b.iterate(6, use_graph=True)
    assert ra.gamma == rb.gamma
    assert np.array_equal(a.image(), b.image())


def test_autotune_picks_a_fixed_order_kernel_at_the_2048_shape():
    n, M = 2048, 4_000_000
    u, v = make_uv(1002, M, n)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 6, 6,
                                                     precision="c128")
    chosen = lt.op.autotune(3)
    assert chosen in (0, 2) and lt.op.get_tuning("grid_kernel") == chosen
    y = torch.from_numpy(utilities.synth_cnormal(7, M)).cuda()
    x1 = lt.op.adjoint_apply(y).cpu().numpy()
    for _ in range(2):
        assert np.array_equal(lt.op.adjoint_apply(y).cpu().numpy(), x1)


def test_tight_frame_proximal_takes_its_own_nu():
    n, M, K = 64, 5000, 12
    u, v, x0, y, eps, orc = problem(1121, n, M)
    ents = [(wavelets.parse_wavelet(a), b) for a, b in DB4]
    xo, st, gam, _ = oracle.padmm(orc, ents, y, eps, K, nu=2.0, l1_nu=1.0)
    xo1, _, _, _ = oracle.padmm(orc, ents, y, eps, K, nu=2.0, l1_nu=2.0)
    assert rel(xo, xo1) > 1e-6  # the two frame bounds give different iterates
    sol, lt, W = solver(u, v, y, n, eps, "c128", orc.norm, K, nu=2.0, l1_proximal_nu=1.0)
    r = sol.iterate(K, use_graph=False)
    assert r.gamma == pytest.approx(gam, rel=1e-12)
    assert rel(sol.image(), xo) < 1e-9
    assert np.allclose(sol.stats(), st, rtol=1e-8, atol=0)


def test_wrong_device_tensors_are_refused_before_the_library():
    n, M = 32, 1000
    u, v = make_uv(1131, M, n)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                     precision="c64")
    op = lt.op
    x = torch.zeros(n * n, dtype=torch.complex64, device="cuda")
    y = torch.zeros(M, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError, match="entries"):
        op.forward(x, out=torch.zeros(M - 1, dtype=torch.complex64, device="cuda"))
    with pytest.raises(TypeError):
        op.forward(x, out=torch.zeros(M, dtype=torch.complex128, device="cuda"))
    with pytest.raises(ValueError, match="contiguous"):
        op.adjoint_apply(y, out=torch.zeros(2 * n * n, dtype=torch.complex64, device="cuda")[::2])
    with pytest.raises(TypeError):
        op.adjoint_apply(y, out=torch.zeros(n * n, dtype=torch.complex64))  # host tensor
    with pytest.raises(ValueError, match="entries"):
        op.normal(x, out=torch.zeros(n * n + 1, dtype=torch.complex64, device="cuda"))
    with pytest.raises(ValueError, match="entries"):
        op.apply_G(x)  # an image where a grid is expected
    with pytest.raises(TypeError):
        op.apply_GH(y.to(torch.complex128))
    W = wavelets.WaveletOperator(DB4, n, n, "c64")
    with pytest.raises(ValueError, match="entries"):
        algorithms.ImagingProximalADMM(op, W, y[:-1], 1.0)
    with pytest.raises(TypeError):
        algorithms.ImagingProximalADMM(op, W, y.to(torch.complex128), 1.0)
    with pytest.raises(ValueError, match="entries"):
        W.analysis(x, out=torch.zeros(n * n - 1, dtype=torch.complex64, device="cuda"))
    # the right tensors still work
    out = torch.empty(M, dtype=torch.complex64, device="cuda")
    assert op.forward(x, out=out) is out
