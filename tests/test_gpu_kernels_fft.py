"""GPU parity of the kernel evaluation and the 2-D FFT.

Kernel values are checked against the reference fixtures of cpp/tests/kernels.cc:35-73 (abs 1e-13
in FP64).  The FFT is checked against the oracle's FFT (itself pinned to a DFT matrix):
FP64 within 1e-12 relative L2, FP32 within 2e-6.
"""
import numpy as np
import pytest
import torch

import oracle
from golden import fixtures
from gpu_helpers import rel
import purify_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["kb", "gauss", "pswf"])
def test_gpu_kernel_fixture_fp64(name):
    x = fixtures.read_reals("kernels", "kernel_x_data")
    expected = fixtures.read_reals("kernels", f"kernel_{name}_data")
    got = purify_amd.kernel_eval(name, x, 6, precision="c128")
    assert np.max(np.abs(got - expected)) < 1e-13


@pytest.mark.parametrize("name", ["kb", "gauss", "pswf"])
def test_gpu_kernel_fixture_fp32(name):
    x = fixtures.read_reals("kernels", "kernel_x_data")
    expected = fixtures.read_reals("kernels", f"kernel_{name}_data")
    got = purify_amd.kernel_eval(name, x, 6, precision="c64")
    assert np.max(np.abs(got - expected)) < 2e-6


@pytest.mark.parametrize("name", ["kb", "gauss", "pswf"])
def test_correction_fixture(name):
    # which=1 evaluates FT(r / ftsize - 0.5); with ftsize = 1 and r = x + 0.5 it is FT(x)
    x = fixtures.read_reals("kernels", "correction_x_data")
    expected = fixtures.read_reals("kernels", f"correction_{name}_data")
    got = purify_amd.kernel_eval(name, x + 0.5, 6, which=1, ftsize=1)
    assert np.max(np.abs(got - expected)) < 1e-13


@pytest.mark.parametrize("kind", ["kb", "kbmin", "gauss_alt", "box"])
@pytest.mark.parametrize("J", [2, 4, 7, 8])
def test_gpu_kernels_vs_oracle(kind, J):
    x = np.linspace(-J / 2, J / 2, 257)
    ref = oracle.kernel_eval(oracle.KERNELS[kind], 0, x, J, 1, 2.0)
    got = purify_amd.kernel_eval(kind, x, J, precision="c128")
    assert np.max(np.abs(got - ref)) < 1e-13


SIZES = [(2, 2), (4, 8), (16, 16), (8, 32), (32, 64), (64, 32), (128, 128), (512, 512), (256, 1024), (1024, 256),
         (2048, 64), (32, 2048)]


@pytest.mark.parametrize("Nv,Nu", SIZES)
@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("sign", [-1, 1])
def test_gpu_fft_vs_oracle(Nv, Nu, precision, sign):
    rng = np.random.default_rng(Nv * 7 + Nu)
    x = rng.normal(size=Nv * Nu) + 1j * rng.normal(size=Nv * Nu)
    ref = oracle.fft2d(Nv, Nu, x, sign)
    dt = torch.complex128 if precision == "c128" else torch.complex64
    got = purify_amd.fft2d(torch.from_numpy(x).to(dt).cuda(), Nv, Nu, sign).cpu().numpy()
    tol = 1e-12 if precision == "c128" else 2e-6
    assert rel(got, ref) < tol


@pytest.mark.parametrize("N", [2048, 4096])
def test_gpu_fft_large_round_trip(N):
    rng = np.random.default_rng(N)
    x = (rng.normal(size=N * N) + 1j * rng.normal(size=N * N)).astype(np.complex64)
    xt = torch.from_numpy(x).cuda()
    X = purify_amd.fft2d(xt, N, N, -1)
    # Parseval for the unitary transform, then the inverse
    assert torch.linalg.vector_norm(X).item() == pytest.approx(torch.linalg.vector_norm(xt).item(), rel=1e-5)
    back = purify_amd.fft2d(X, N, N, +1).cpu().numpy()
    assert rel(back, x) < 3e-6
    # a plane wave lands on one bin
    k0, l0 = 17, N // 3
    p, q = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    w = np.exp(2j * np.pi * (k0 * p + l0 * q) / N).ravel()
    W = purify_amd.fft2d(torch.from_numpy(w).cuda(), N, N, -1).cpu().numpy().reshape(N, N)
    assert abs(W[k0, l0] - N) < 1e-8 * N
    W[k0, l0] = 0
    assert np.max(np.abs(W)) < 1e-9 * N
