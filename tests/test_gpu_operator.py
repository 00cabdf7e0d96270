"""GPU parity of the measurement operator Phi / Phi^dagger against the oracle and the reference fixtures.

Tolerances (relative L2 against the FP64 oracle on the same inputs): complex128 path 1e-12,
complex64 path 1e-5.  Reference fixtures: cpp/tests/operators.cc:65-92 (G, G^H, 1e-5) and
cpp/tests/measurement_operator.cc:14-84 (full Phi on M31, max-normalised, 1e-6).
"""
import numpy as np
import pytest
import torch

import oracle
from golden import fixtures
from gpu_helpers import KIND, isapprox, make_uv, pair, rel
from purify_amd import kernel, measurementoperator, utilities, vis_params, vis_units
from purify_amd._lib import PurifyError

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}
DT = {"c128": torch.complex128, "c64": torch.complex64}


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_gridding_fixture(precision):
    u = fixtures.read_reals("operators", "u_data")
    v = fixtures.read_reals("operators", "v_data")
    op = measurementoperator.init_degrid_operator_2d(u, v, None, None, 16, 16, 2.0, 0, 1e-4, kernel.kb, 4, 4,
                                                     precision=precision).op
    x = torch.from_numpy(fixtures.read_complex("operators", "direct_input_data")).to(DT[precision]).cuda()
    y = op.apply_G(x).cpu().numpy()
    assert isapprox(y, fixtures.read_complex("operators", "expected_direct_data"), 1e-5)
    yin = torch.from_numpy(fixtures.read_complex("operators", "indirect_input_data")).to(DT[precision]).cuda()
    g = op.apply_GH(yin).cpu().numpy()
    assert isapprox(g, fixtures.read_complex("operators", "expected_indirect_data"), 1e-5)
    # and against the oracle's explicit CSR
    G = oracle.gridding_matrix(u, v, np.ones(10, np.complex128), 16, 16, 2.0, oracle.KB, 4, 4, 4, 4)
    assert rel(y, G.dot(x.cpu().numpy().astype(np.complex128))) < TOL[precision]
    assert rel(g, G.adjoint().dot(yin.cpu().numpy().astype(np.complex128))) < TOL[precision]


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("name,J", [("kb", 4), ("pswf", 6), ("gauss", 6)])
def test_m31_fixture(precision, name, J):
    u = fixtures.read_reals("measurement_operator", "u")
    v = fixtures.read_reals("measurement_operator", "v")
    img, _ = fixtures.read_fits_image("images", "M31.fits")
    M = u.size
    # the reference test passes (v, u) into the (u, v) slots, power_iters = 1
    lt = measurementoperator.init_degrid_operator_2d(v, u, np.zeros(M), np.ones(M), 256, 256, 2.0, 1, 1e-9, name, J, J,
                                                     precision=precision)
    y = lt * img
    yexp = fixtures.read_complex("measurement_operator", f"{name}_exact_data")
    assert isapprox(y / np.max(np.abs(y)), yexp / np.max(np.abs(yexp)), 1e-6)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_c1_shape_vs_oracle(precision):
    # 256^2 image, oversample 2, KB J = 4, M = 1e4, uv ~ N(0, pi/3) rad
    n, M = 256, 10_000
    u, v = make_uv(1000, M, n)
    gpu, orc = pair(u, v, n, n, precision=precision)
    x = utilities.synth_image(1000, n, n, 20)
    y = gpu * x
    yo = orc.forward(x)
    assert rel(y, yo) < TOL[precision]
    r = utilities.synth_cnormal(77, M)
    xa = gpu.adjoint() * r
    xo = orc.adjoint(r)
    assert rel(xa, xo) < TOL[precision]


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_c2_subset_vs_oracle(precision):
    # 1024^2 image, J = 6, a 2e5 subset of the 2e6-visibility precision study
    n, M = 1024, 200_000
    u, v = make_uv(1002, M, n)
    gpu, orc = pair(u, v, n, n, Ju=6, Jv=6, precision=precision)
    x = utilities.synth_image(1002, n, n, 20)
    assert rel(gpu * x, orc.forward(x)) < TOL[precision]
    r = utilities.synth_cnormal(1002, M)
    assert rel(gpu.adjoint() * r, orc.adjoint(r)) < TOL[precision]


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_adjoint_dot_product(precision):
    n, M = 512, 50_000
    u, v = make_uv(5, M, n)
    w = utilities.synth_cnormal(6, M)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, w, n, n, 2.0, 0, 1e-4, "kb", 6, 6,
                                                     precision=precision)
    x = utilities.synth_cnormal(8, n * n)
    y = utilities.synth_cnormal(9, M)
    lhs = np.vdot(y, lt * x)
    rhs = np.vdot(lt.adjoint() * y, x)
    assert abs(lhs - rhs) / abs(lhs) < TOL[precision]


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_gridding_is_deterministic(precision):
    n, M = 512, 300_000
    u, v = make_uv(11, M, n, sigma=0.4)  # dense centre: many visibilities share cells
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 6, 6,
                                                     precision=precision)
    y = torch.from_numpy(utilities.synth_cnormal(12, M)).to(DT[precision]).cuda()
    a = lt.op.apply_GH(y)
    b = lt.op.apply_GH(y)
    c = lt.op.adjoint_apply(y)
    d = lt.op.adjoint_apply(y)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(c, d)


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_gridding_variants_agree_and_are_deterministic(precision, variant):
    n, M = 256, 60_000
    u, v = make_uv(13, M, n, sigma=0.5)
    w = utilities.synth_cnormal(14, M)
    gpu, orc = pair(u, v, n, n, weights=w, Ju=6, Jv=6, precision=precision)
    gpu.op.set_tuning("grid_kernel", variant)
    assert gpu.op.get_tuning("grid_kernel") == variant
    y = utilities.synth_cnormal(15, M)
    a = gpu.adjoint() * y
    assert rel(a, orc.adjoint(y)) < TOL[precision]
    b = gpu.adjoint() * y
    assert np.array_equal(a, b)
    chosen = gpu.op.autotune(2)
    assert chosen in (0, 1, 2) and gpu.op.get_tuning("grid_kernel") == chosen
    assert gpu.op.get_tuning("fft_columns") in (0, 1)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < TOL[precision]


def test_matrix_core_gridding_needs_room_for_a_tile():
    # 8 x 8 image -> 16 x 16 grid: smaller than 16 + J - 1, so only the other variants exist
    u, v = make_uv(19, 200, 8)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, 8, 8, 2.0, 0, 1e-4, "kb", 4, 4,
                                                     precision="c128")
    assert lt.op.get_tuning("grid_kernel") != 2
    with pytest.raises(PurifyError, match="matrix-core"):
        lt.op.set_tuning("grid_kernel", 2)


@pytest.mark.parametrize("kern,J", [("kb", 6), ("kb", 4), ("gauss", 6), ("pswf", 6), ("kbmin", 8)])
@pytest.mark.parametrize("table", [2, 1, 0])
def test_complex64_tap_table_and_exact_kernels(kern, J, table):
    n, M = 512, 100_000
    u, v = make_uv(16 + J, M, n)
    gpu, orc = pair(u, v, n, n, kernel=kern, Ju=J, Jv=J, precision="c64")
    assert gpu.op.get_tuning("kernel_table") == 1
    gpu.op.set_tuning("kernel_table", table)
    x = utilities.synth_image(17, n, n, 20)
    assert rel(gpu * x, orc.forward(x)) < 1e-5
    y = utilities.synth_cnormal(18, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < 1e-5


@pytest.mark.parametrize("kern,J", [("gauss", 6), ("pswf", 6), ("box", 3), ("kbmin", 5), ("gauss_alt", 4),
                                    ("kb", 2), ("kb", 3), ("kb", 5), ("kb", 7), ("kb", 8), ("kb", 12)])
def test_kernels_and_supports_vs_oracle(kern, J):
    n, M = 64, 3000
    u, v = make_uv(21 + J, M, n)
    gpu, orc = pair(u, v, n, n, kernel=kern, Ju=J, Jv=J)
    x = utilities.synth_cnormal(3, n * n)
    assert rel(gpu * x, orc.forward(x)) < 1e-12
    y = utilities.synth_cnormal(4, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < 1e-12


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_unequal_supports_nonsquare_weights_and_wrap(precision):
    ny, nx, M = 32, 128, 4000  # non-square: 64 x 256 grid
    rng = np.random.default_rng(31)
    # coordinates up to and beyond the grid edge to exercise the circular wrap
    u = rng.uniform(-1.2, 1.2, M) * nx
    v = rng.uniform(-1.2, 1.2, M) * ny
    w = utilities.synth_cnormal(32, M)
    # Ju != Jv: the Gaussian kernel is defined everywhere, so the reference's kernel-slot swap
    # (operators.h:322-324) gives finite values that both sides must reproduce
    gpu, orc = pair(u, v, ny, nx, weights=w, kernel="gauss", Ju=4, Jv=6, precision=precision, tile=(16, 8))
    x = utilities.synth_cnormal(33, nx * ny)
    assert rel(gpu * x, orc.forward(x)) < TOL[precision]
    y = utilities.synth_cnormal(34, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < TOL[precision]
    # same geometry with equal KB supports
    gpu, orc = pair(u, v, ny, nx, weights=w, kernel="kb", Ju=6, Jv=6, precision=precision, tile=(16, 8))
    assert rel(gpu * x, orc.forward(x)) < TOL[precision]
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < TOL[precision]


def test_kb_unequal_supports_is_nan_like_the_reference():
    # With Ju < Jv the v taps reach |x| > Ju/2 of the u-support KB kernel in the v slot, where the
    # reference evaluates sqrt of a negative number (kernels.cc:20): every visibility is NaN.
    n, M = 32, 50
    u, v = make_uv(35, M, n)
    gpu, orc = pair(u, v, n, n, kernel="kb", Ju=4, Jv=6)
    x = utilities.synth_cnormal(36, n * n)
    assert np.all(np.isnan(orc.forward(x)))
    assert np.all(np.isnan(gpu * x))


def test_single_and_empty_visibility_sets():
    n = 16
    gpu, orc = pair(np.array([3.3]), np.array([-7.9]), n, n, precision="c128")
    x = utilities.synth_cnormal(1, n * n)
    assert rel(gpu * x, orc.forward(x)) < 1e-12
    assert rel(gpu.adjoint() * np.array([1 + 2j]), orc.adjoint(np.array([1 + 2j]))) < 1e-12
    empty = measurementoperator.init_degrid_operator_2d(np.zeros(0), np.zeros(0), None, None, n, n, 2.0, 0, 1e-4,
                                                        precision="c128")
    assert (empty * x).size == 0
    assert np.all(empty.adjoint() * np.zeros(0, np.complex128) == 0)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_power_method_and_normalisation(precision):
    n, M = 128, 5000
    u, v = make_uv(41, M, n)
    orc = oracle.Operator(u, v, None, None, n, n, 2.0, oracle.KB, 4, 4)
    ref = orc.power_method(100, 1e-4)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 100, 1e-4, "kb", 4, 4,
                                                     precision=precision)
    assert lt.norm == pytest.approx(ref, rel=1e-10 if precision == "c128" else 1e-5)
    # the normalised operator has norm 1 +- tol (cpp/tests/operators.cc:147-154)
    assert abs(lt.op.power_method(1000, 1e-6) - 1.0) < 1e-5
    # and it is Phi / norm
    orc.set_norm(lt.norm)
    x = utilities.synth_cnormal(42, n * n)
    assert rel(lt * x, orc.forward(x)) < TOL[precision]


def test_device_tensor_apply_matches_host_apply():
    n, M = 256, 20_000
    u, v = make_uv(51, M, n)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 10, 1e-4, precision="c64")
    x = utilities.synth_image(51, n, n, 20)
    yh = lt * x
    yd = lt.op.forward(torch.from_numpy(x).to(torch.complex64).cuda()).cpu().numpy()
    assert rel(yd, yh) < 1e-7
    xd = lt.op.adjoint_apply(torch.from_numpy(yh).to(torch.complex64).cuda()).cpu().numpy()
    assert rel(xd, lt.adjoint() * yh) < 1e-7


def test_vis_params_overload_units():
    # lambda -> radians -> pixels conversion of operators.h:405-410
    n, M, cell = 128, 2000, 3.0
    u_rad, v_rad = utilities.synth_uv(61, M, 0.5)
    scale = 180 * 3600 / cell / np.pi
    uvl = vis_params(u_rad / (2 * np.pi) * scale * 2 * np.pi, v_rad / (2 * np.pi) * scale * 2 * np.pi,
                     units=vis_units.lambda_)
    lt = measurementoperator.init_degrid_operator_2d(uvl, n, n, cell, cell, 2.0, 0, 1e-4, precision="c128")
    up, vp = oracle.uv_to_pixels(uvl.u, uvl.v, 0, cell, cell, n, n, 2.0)
    orc = oracle.Operator(up, vp, None, None, n, n, 2.0, oracle.KB, 4, 4)
    x = utilities.synth_cnormal(62, n * n)
    assert rel(lt * x, orc.forward(x)) < 1e-12


def test_vis_params_overload_radians_units_take_uv_scale_only():
    # operators.h:408-410: radians -> uv_scale (utilities.cc:258-278) -> pixels, u / (2 pi) * floor(os n)
    n, M = 64, 1500
    u_rad, v_rad = utilities.synth_uv(63, M, 0.8)
    uvr = vis_params(u_rad, v_rad, units=vis_units.radians)
    lt = measurementoperator.init_degrid_operator_2d(uvr, n, n, 1.0, 1.0, 2.0, 0, 1e-4, precision="c128")
    orc = oracle.Operator(utilities.radians_to_pixels(u_rad, n), utilities.radians_to_pixels(v_rad, n), None, None,
                          n, n, 2.0, oracle.KB, 4, 4)
    x = utilities.synth_cnormal(64, n * n)
    assert rel(lt * x, orc.forward(x)) < 1e-12
    up, vp = oracle.uv_to_pixels(u_rad, v_rad, 1, 1.0, 1.0, n, n, 2.0)
    assert np.allclose(up, u_rad / (2 * np.pi) * 2 * n) and np.allclose(vp, v_rad / (2 * np.pi) * 2 * n)


def test_construction_errors_carry_reference_messages():
    u = np.zeros(4)
    with pytest.raises(PurifyError, match="PSWF requires a support of 6"):
        measurementoperator.init_degrid_operator_2d(u, u, None, None, 16, 16, 2.0, 0, 1e-4, "pswf", 4, 4)
    with pytest.raises(PurifyError, match="powers of two"):
        measurementoperator.init_degrid_operator_2d(u, u, None, None, 10, 10, 2.0, 0, 1e-4)
    with pytest.raises(RuntimeError, match="Size of u and v"):
        measurementoperator.init_degrid_operator_2d(u, np.zeros(3), None, None, 16, 16)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_fft_column_modes_agree(precision):
    # 1024^2 image -> 2048^2 grid: one-pass columns (default) against the N1 x N2 two-step split
    n, M = 1024, 50_000
    u, v = make_uv(71, M, n)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 6, 6,
                                                     precision=precision)
    assert lt.op.get_tuning("fft_columns") == 0
    x = utilities.synth_image(72, n, n, 20)
    y = utilities.synth_cnormal(73, M)
    a, b = lt * x, lt.adjoint() * y
    lt.op.set_tuning("fft_columns", 1)
    assert lt.op.get_tuning("fft_columns") == 1
    c, d = lt * x, lt.adjoint() * y
    tol = 1e-13 if precision == "c128" else 2e-6
    assert rel(a, c) < tol and rel(b, d) < tol
    with pytest.raises(PurifyError):
        lt.op.set_tuning("fft_columns", 2)


@pytest.mark.parametrize("precision,M", [("c64", 16384), ("c64", 16385), ("c64", 49153), ("c64", 63),
                                         ("c128", 8192), ("c128", 8193), ("c128", 24577), ("c128", 1)])
def test_permutation_chunk_and_bucket_boundaries(precision, M):
    # the caller <-> sorted permutation runs in chunks and buckets of 16384 (c64) / 8192 (c128)
    # visibilities: exact multiples, one past, a ragged last chunk and a single partial chunk
    n = 64
    u, v = make_uv(81 + M, M, n)
    w = utilities.synth_cnormal(82, M)
    gpu, orc = pair(u, v, n, n, weights=w, precision=precision)
    x = utilities.synth_image(83, n, n, 10)
    assert rel(gpu * x, orc.forward(x)) < TOL[precision]
    y = utilities.synth_cnormal(84, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < TOL[precision]


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_row_pass_grid_sizes_give_identical_results(precision):
    # each line's arithmetic does not depend on which block transforms it or on the prefetch
    n, M = 512, 20_000
    u, v = make_uv(91, M, n)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, n, n, 2.0, 0, 1e-4, "kb", 4, 4,
                                                     precision=precision)
    assert lt.op.get_tuning("fft_row_blocks") == (3 if precision == "c64" else 2)
    x = utilities.synth_image(92, n, n, 20)
    y = utilities.synth_cnormal(93, M)
    ref = (lt * x, lt.adjoint() * y)
    for rb in (0, 1, 3, 64):
        lt.op.set_tuning("fft_row_blocks", rb)
        assert lt.op.get_tuning("fft_row_blocks") == rb
        assert np.array_equal(lt * x, ref[0]) and np.array_equal(lt.adjoint() * y, ref[1])
    with pytest.raises(PurifyError):
        lt.op.set_tuning("fft_row_blocks", 65)
    # the autotuner times the row pass's grids on whole pairs and keeps one of its candidates
    lt.op.autotune(1)
    assert lt.op.get_tuning("fft_row_blocks") in (0, 2, 3, 4)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_matrix_core_wave_counts_give_identical_results(precision):
    # grid_waves W: wave w grids tiles w, w + W, ... with its batches pipelined across tile boundaries;
    # every tile's sum runs in the same order whatever W is
    n, M = 128, 30_000
    u, v = make_uv(95, M, n, sigma=0.5)
    w = utilities.synth_cnormal(96, M)
    gpu, orc = pair(u, v, n, n, weights=w, Ju=6, Jv=6, precision=precision)
    assert gpu.op.get_tuning("grid_kernel") == 2 and gpu.op.get_tuning("grid_waves") == 0
    y = utilities.synth_cnormal(97, M)
    ref = gpu.adjoint() * y
    assert rel(ref, orc.adjoint(y)) < TOL[precision]
    for gw in (1, 7, 64, 10**6):
        gpu.op.set_tuning("grid_waves", gw)
        assert np.array_equal(gpu.adjoint() * y, ref)
    with pytest.raises(PurifyError):
        gpu.op.set_tuning("grid_waves", -1)


@pytest.mark.parametrize("kernel,J", [("kb", 4), ("kb", 6), ("kb", 8), ("gauss", 6), ("pswf", 6), ("kbmin", 5)])
def test_tap_polynomials_vs_oracle_and_direct_evaluation(kernel, J):
    # complex64 with Ju = Jv: per-tap polynomials in the fractional offset (degree 9, checked on the
    # host against the exact kernel at build) replace the per-tap kernel evaluation
    n, M = 128, 20_000
    u, v = make_uv(99 + J, M, n, sigma=0.6)
    w = utilities.synth_cnormal(98, M)
    gpu, orc = pair(u, v, n, n, weights=w, kernel=kernel, Ju=J, Jv=J, precision="c64")
    assert gpu.op.get_tuning("kernel_poly") == 1
    x = utilities.synth_image(97, n, n, 20)
    y = utilities.synth_cnormal(96, M)
    a, b = gpu * x, gpu.adjoint() * y
    assert rel(a, orc.forward(x)) < 1e-5 and rel(b, orc.adjoint(y)) < 1e-5
    gpu.op.set_tuning("kernel_poly", 0)
    assert gpu.op.get_tuning("kernel_poly") == 0
    assert rel(gpu * x, a) < 2e-6 and rel(gpu.adjoint() * y, b) < 2e-6


def test_tap_polynomials_only_where_they_apply():
    u, v = make_uv(95, 500, 32)
    lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, 32, 32, 2.0, 0, 1e-4, "kb", 4, 4,
                                                     precision="c128")
    assert lt.op.get_tuning("kernel_poly") == 0
    with pytest.raises(PurifyError, match="polynomials"):
        lt.op.set_tuning("kernel_poly", 1)
    # unequal supports (the kernel-slot swap puts taps outside the support) and the box kernel: none
    for kern, Ju, Jv in (("kb", 4, 6), ("box", 4, 4)):
        lt = measurementoperator.init_degrid_operator_2d(u, v, None, None, 32, 32, 2.0, 0, 1e-4, kern, Ju, Jv,
                                                         precision="c64")
        assert lt.op.get_tuning("kernel_poly") == 0


@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("w_term", [False, True])
def test_tile_shape_chosen_by_timing_equals_the_explicit_build(precision, w_term):
    # tile="auto": one build per candidate tile shape (and complex64 permutation bucket), the fastest
    # Phi + Phi^dagger pair kept.  The tile shape sets the sort order and the bucket only moves data, so
    # the tuned operator equals the one built with the chosen shape, bit for bit, normalisation included.
    n, M = 256, 100_000
    u, v = make_uv(121, M, n)
    w = np.random.default_rng(122).uniform(0, 150, M) if w_term else None
    wt = utilities.synth_cnormal(123, M)
    kw = dict(precision=precision, w_term=w_term, Jw=4, cellx=3600.0 / n, celly=3600.0 / n)
    args = (u, v, w, wt, n, n, 2.0, 20, 1e-4, "kb", 6, 6)
    lta = measurementoperator.init_degrid_operator_2d(*args, tile="auto", **kw)
    tu, tv = lta.op.get_tuning("tile_u"), lta.op.get_tuning("tile_v")
    assert (tu, tv) in ((32, 32), (64, 32), (32, 64), (64, 64))
    bucket = lta.op.get_tuning("perm_bucket")
    if precision == "c64":
        assert bucket in (8192, 16384)
    else:
        assert bucket == 0
    lte = measurementoperator.init_degrid_operator_2d(*args, tile=(tu, tv), **kw)
    x = utilities.synth_image(124, n, n, 20)
    y = utilities.synth_cnormal(125, M)
    assert np.array_equal(lta * x, lte * x)
    assert np.array_equal(lta.adjoint() * y, lte.adjoint() * y)
    orc = oracle.Operator(u, v, w, wt, n, n, 2.0, oracle.KB, 6, 6, w_term, 4, 3600.0 / n, 3600.0 / n)
    nrm = lta.op.norm
    assert rel(lta * x, orc.forward(x) / nrm) < TOL[precision] * (10 if w_term else 1)
    with pytest.raises(ValueError):
        measurementoperator.init_degrid_operator_2d(*args, tile="fast", **kw)
