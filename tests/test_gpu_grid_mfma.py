"""GPU parity of the matrix-core gridding (grid_kernel = 2) at its edges.

G^H (the reference's sparse adjoint, cpp/purify/operators.h:123-145, utilities.h:144-159) summed per
16 x 16 tile as rank-4 updates.  Every case is checked against the FP64 oracle (relative L2 <= 1e-5
for complex64, 1e-12 for complex128), and results must be bit-identical run to run and for every wave
count, with complex weights in both tap modes.
"""
import numpy as np
import pytest
import torch

from gpu_helpers import make_uv, pair, rel
from purify_amd import utilities

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


@pytest.mark.parametrize("kern,Ju,Jv", [("kb", 2, 2), ("kb", 3, 3), ("kb", 5, 5), ("kb", 7, 7), ("kb", 8, 8),
                                        ("gauss", 9, 9), ("gauss", 12, 12), ("gauss", 16, 16), ("gauss", 4, 6)])
@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_matrix_core_gridding_every_support(kern, Ju, Jv, precision):
    # supports 2..8 run the compile-time kernels, 9..16 and unequal supports the run-time one
    n, M = 128, 40_000
    u, v = make_uv(40 + Ju + 3 * Jv, M, n, sigma=0.7)
    w = utilities.synth_cnormal(41, M)
    gpu, orc = pair(u, v, n, n, weights=w, kernel=kern, Ju=Ju, Jv=Jv, precision=precision)
    assert gpu.op.get_tuning("grid_kernel") == 2
    y = utilities.synth_cnormal(42, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < TOL[precision]


def test_matrix_core_gridding_nonsquare_grid_and_edge_wrap():
    # uniform uv over the whole band: many footprints wrap around the grid edges and meet tiles of
    # both sides; 64 x 32 image -> 64 x 128 grid
    ny, nx, M = 32, 64, 30_000
    rng = np.random.default_rng(61)
    u = utilities.radians_to_pixels(rng.uniform(-np.pi, np.pi, M), nx)
    v = utilities.radians_to_pixels(rng.uniform(-np.pi, np.pi, M), ny)
    w = utilities.synth_cnormal(62, M)
    gpu, orc = pair(u, v, ny, nx, weights=w, Ju=5, Jv=5, precision="c64")
    assert gpu.op.get_tuning("grid_kernel") == 2
    y = utilities.synth_cnormal(63, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < 1e-5


@pytest.mark.parametrize("M", [1, 15, 16, 17, 63, 64, 65, 130, 1000])
def test_matrix_core_gridding_coincident_visibilities_across_batch_boundaries(M):
    # every visibility on one point: one tile's list holds them all, in batches of 64 and MFMA steps of
    # four entries; the compensated tile sums fold every four batches
    n = 64
    u = np.full(M, 3.3)
    v = np.full(M, -7.6)
    gpu, orc = pair(u, v, n, n, Ju=6, Jv=6, precision="c64")
    y = utilities.synth_cnormal(70 + M, M)
    assert rel(gpu.adjoint() * y, orc.adjoint(y)) < 1e-5


@pytest.mark.parametrize("poly", [1, 0])
def test_matrix_core_gridding_wave_count_independent_with_complex_weights(poly):
    # complex weights enter each entry through a complex product whose fused multiply-adds are explicit,
    # so every inlined copy of the batch code rounds alike; dense centre: tiles with thousands of entries
    n, M = 256, 300_000
    u, v = make_uv(81, M, n, sigma=0.4)
    w = utilities.synth_cnormal(82, M)
    gpu, orc = pair(u, v, n, n, weights=w, Ju=6, Jv=6, precision="c64")
    gpu.op.set_tuning("kernel_poly", poly)
    gpu.op.set_tuning("kernel_table", 0)
    y = utilities.synth_cnormal(83, M)
    ref = gpu.adjoint() * y
    assert rel(ref, orc.adjoint(y)) < 1e-5
    assert np.array_equal(gpu.adjoint() * y, ref)
    yd = torch.from_numpy(y).to(torch.complex64).cuda()
    g1 = gpu.op.apply_GH(yd)
    for gw in (1, 7, 64, 10**6):
        gpu.op.set_tuning("grid_waves", gw)
        assert np.array_equal(gpu.adjoint() * y, ref)
        assert torch.equal(gpu.op.apply_GH(yd), g1)
    gpu.op.set_tuning("grid_waves", 0)


def test_matrix_core_gridding_in_the_compact_operator_and_dot_test():
    # Phi^dagger Phi runs degridding and gridding in the sorted order; the dot test <y, Phi x> =
    # <Phi^dagger y, x> holds to the complex64 bar
    n, M = 256, 80_000
    u, v = make_uv(91, M, n)
    gpu, orc = pair(u, v, n, n, Ju=6, Jv=6, precision="c64")
    x = utilities.synth_image(92, n, n, 20).astype(np.complex128)
    y = utilities.synth_cnormal(93, M)
    lhs = np.vdot(y, gpu * x)
    rhs = np.vdot(gpu.adjoint() * y, x)
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    assert rel(gpu.op.normal(x), orc.adjoint(orc.forward(x))) < 1e-5
