"""GPU parity of the w-projection measurement operator (w_term = true) against the oracle.

The operator's rows are the projection kernels of cpp/purify/projection_kernels.cc:59-85 placed by
cpp/purify/operators.cc:49-99 (G receives w - mean(w), operators.h:323).  Tolerances: complex128 1e-12,
complex64 1e-5 relative L2 against the FP64 oracle.
"""
import numpy as np
import pytest
import torch

import oracle
from gpu_helpers import KIND, make_uv, rel
from purify_amd import measurementoperator, utilities, vis_params, vis_units
from purify_amd._lib import PurifyError

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def wpair(u, v, w, ny, nx, weights=None, kernel="kb", Ju=4, Jv=4, Jw=6, cell=None, precision="c128"):
    # field of view 1 degree: cell = 3600 / n arcsec (the w-projection shape's convention)
    cellx = cell if cell is not None else 3600.0 / nx
    celly = cell if cell is not None else 3600.0 / ny
    gpu = measurementoperator.init_degrid_operator_2d(u, v, w, weights, ny, nx, 2.0, 0, 1e-4, kernel, Ju, Jv, True,
                                                      Jw, cellx, celly, precision=precision)
    orc = oracle.Operator(u, v, w, weights, ny, nx, 2.0, KIND[kernel], Ju, Jv, True, Jw, cellx, celly)
    return gpu, orc


def w_uniform(seed, M, wmax=200.0):
    return np.random.default_rng(seed).uniform(0.0, wmax, M)


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("J,Jw", [(4, 6), (6, 6), (4, 3)])
def test_wprojection_operator_vs_oracle(precision, J, Jw):
    n, M = 64, 3000
    u, v = make_uv(301 + J, M, n)
    w = w_uniform(302, M)
    gpu, orc = wpair(u, v, w, n, n, Ju=J, Jv=J, Jw=Jw, precision=precision)
    info = gpu.op.info()
    assert info.w_term == 1 and info.nnz_w == M * (J + Jw - 1) ** 2 == orc.nnz
    x = utilities.synth_image(303, n, n, 20)
    assert rel(gpu * x, orc.forward(x)) < TOL[precision]
    y = utilities.synth_cnormal(304, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < TOL[precision]


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_wprojection_weights_nonsquare_wrap_and_unequal_supports(precision):
    ny, nx, M = 32, 64, 2500
    rng = np.random.default_rng(311)
    u = rng.uniform(-1.1, 1.1, M) * nx  # beyond the grid edge: circular wrap of the rows
    v = rng.uniform(-1.1, 1.1, M) * ny
    w = rng.uniform(0, 150, M)
    wt = utilities.synth_cnormal(312, M)
    # Gaussian kernel: finite for Ju != Jv despite the kernel-slot swap
    gpu, orc = wpair(u, v, w, ny, nx, weights=wt, kernel="gauss", Ju=4, Jv=6, Jw=5, precision=precision)
    x = utilities.synth_cnormal(313, nx * ny)
    assert rel(gpu * x, orc.forward(x)) < TOL[precision]
    y = utilities.synth_cnormal(314, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < TOL[precision]


def test_wprojection_with_w_zero_and_jw_one_is_the_plain_operator():
    # w - mean(w) = 0 makes every chirp sample take the reference's "else 1" branch
    # (projection_kernels.cc:40-45), so with Jw = 1 each row is the plain Kaiser-Bessel row.
    n, M = 64, 4000
    u, v = make_uv(321, M, n)
    w = np.full(M, 37.5)
    wt = utilities.synth_cnormal(322, M)
    lw = measurementoperator.init_degrid_operator_2d(u, v, w, wt, n, n, 2.0, 0, 1e-4, "kb", 6, 6, True, 1, 30.0, 30.0,
                                                     precision="c128")
    l0 = measurementoperator.init_degrid_operator_2d(u, v, None, wt, n, n, 2.0, 0, 1e-4, "kb", 6, 6,
                                                     precision="c128")
    x = utilities.synth_cnormal(323, n * n)
    assert rel(lw * x, l0 * x) < 1e-13
    y = utilities.synth_cnormal(324, M)
    assert rel(lw.adjoint() * y, l0.adjoint() * y) < 1e-13


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_wprojection_adjoint_dot_product_and_determinism(precision):
    n, M = 128, 20_000
    u, v = make_uv(331, M, n, sigma=0.6)
    w = w_uniform(332, M)
    lt = measurementoperator.init_degrid_operator_2d(u, v, w, None, n, n, 2.0, 0, 1e-4, "kb", 6, 6, True, 6,
                                                     3600.0 / n, 3600.0 / n, precision=precision)
    x = utilities.synth_cnormal(333, n * n)
    y = utilities.synth_cnormal(334, M)
    lhs = np.vdot(y, lt * x)
    rhs = np.vdot(lt.adjoint() * y, x)
    assert abs(lhs - rhs) / abs(lhs) < TOL[precision]
    dt = torch.complex128 if precision == "c128" else torch.complex64
    yd = torch.from_numpy(y).to(dt).cuda()
    a = lt.op.adjoint_apply(yd)
    b = lt.op.adjoint_apply(yd)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    # autotuning keeps the one gridding kernel of projection rows and picks a column FFT mode
    g0 = lt.op.get_tuning("grid_kernel")
    assert lt.op.autotune(2) == g0
    c = lt.op.adjoint_apply(yd)
    torch.cuda.synchronize()
    assert (c - a).abs().max().item() <= TOL[precision] * a.abs().max().item()


def test_wprojection_power_method_and_vis_params_units():
    n, M, cell = 64, 1500, 60.0
    u_rad, v_rad = utilities.synth_uv(341, M, 0.8)
    scale = 180 * 3600 / cell / np.pi
    w = w_uniform(342, M, 100.0)
    uvl = vis_params(u_rad * scale, v_rad * scale, w=w, units=vis_units.lambda_)
    lt = measurementoperator.init_degrid_operator_2d(uvl, n, n, cell, cell, 2.0, 50, 1e-6, "kb", 4, 4, True, 4,
                                                     precision="c128")
    up, vp = oracle.uv_to_pixels(uvl.u, uvl.v, 0, cell, cell, n, n, 2.0)
    orc = oracle.Operator(up, vp, w, None, n, n, 2.0, oracle.KB, 4, 4, True, 4, cell, cell, power_iters=50,
                          power_tol=1e-6)
    assert lt.norm == pytest.approx(orc.norm, rel=1e-10)
    x = utilities.synth_cnormal(343, n * n)
    assert rel(lt * x, orc.forward(x)) < 1e-12


def test_wprojection_construction_errors():
    u = np.zeros(4)
    with pytest.raises(PurifyError, match="wider than the grid"):
        # 4 x 4 image -> 8 x 8 grid, rows of 6 + 6 - 1 = 11 taps
        measurementoperator.init_degrid_operator_2d(u, u, u, None, 4, 4, 2.0, 0, 1e-4, "kb", 6, 6, True, 6)
    with pytest.raises(RuntimeError, match="Size of u and w"):
        measurementoperator.init_degrid_operator_2d(u, u, np.zeros(3), None, 16, 16, 2.0, 0, 1e-4, "kb", 4, 4, True, 6)


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("tile", [None, (64, 32), (32, 16)])
def test_owner_tile_gridding_against_the_tile_lists(precision, tile):
    # 11 x 11 rows: each row summed once into its owner tile extended by K - 1, the extensions combined
    # in a fixed order; the tile-list walk sums the same products in another order.  Visibilities crowd
    # the centre (owner tiles cut into several slots) and reach the grid's edges (extensions wrap)
    from purify_amd._lib import PurifyError
    n, M = 64, 20_000
    u, v = make_uv(341, M, n, sigma=0.6)
    u[:4] = utilities.radians_to_pixels(np.array([np.pi * 0.999, -np.pi, 0.0, np.pi * 0.5]), n, 2.0)
    v[:4] = utilities.radians_to_pixels(np.array([-np.pi, np.pi * 0.999, np.pi * 0.999, -np.pi * 0.999]), n, 2.0)
    w = w_uniform(342, M)
    wt = utilities.synth_cnormal(343, M) + 2.0
    kw = {"tile": tile} if tile else {}
    lt = measurementoperator.init_degrid_operator_2d(u, v, w, wt, n, n, 2.0, 0, 1e-4, "kb", 6, 6, True, 6,
                                                     3600.0 / n, 3600.0 / n, precision=precision, **kw)
    if tile:
        assert (lt.op.get_tuning("tile_u"), lt.op.get_tuning("tile_v")) == tile
    assert lt.op.get_tuning("w_footprint") == 11 and lt.op.get_tuning("w_grid_owner") == 1
    y = utilities.synth_cnormal(344, M)
    a = lt.adjoint() * y
    assert np.array_equal(a, lt.adjoint() * y)
    lt.op.set_tuning("w_grid_owner", 0)
    assert lt.op.get_tuning("w_grid_owner") == 0
    b = lt.adjoint() * y
    assert rel(a, b) < TOL[precision]
    lt.op.set_tuning("w_grid_owner", 1)
    assert np.array_equal(a, lt.adjoint() * y)
    with pytest.raises(PurifyError, match="0 or 1"):
        lt.op.set_tuning("w_grid_owner", 2)
    plain = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                        precision=precision)
    assert plain.op.get_tuning("w_grid_owner") == 0
    with pytest.raises(PurifyError, match="owner-tile"):
        plain.op.set_tuning("w_grid_owner", 1)
