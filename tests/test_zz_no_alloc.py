"""pf_fft2d_noalloc (the unitary 2-D FFT that allocates nothing: its twiddle table and workspace live in a
static device array) against pf_fft2d on the same inputs, through the C-ABI: two calls in a row, the
second on a smaller grid (the array keeps the first call's contents), both directions and precisions,
in place, and the size limit.  Tolerance: the FP64 bar of the operator (1e-12 relative L2) for
complex128, the FP32 bar (1e-5) for complex64; the twiddles differ only in their last bit."""
import ctypes as C

import numpy as np
import pytest
import torch

from purify_amd import _lib

pytestmark = pytest.mark.gpu

LIMIT = 32 << 20  # PF_FFT2D_NOALLOC_MAX_BYTES


def fft(name, Nv, Nu, prec, sign, x, out):
    s = torch.cuda.current_stream().cuda_stream
    f = getattr(_lib.lib(), name)
    if name == "pf_fft2d":
        st = f(Nv, Nu, prec, sign, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), None, C.c_void_p(s))
    else:
        st = f(Nv, Nu, prec, sign, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), C.c_void_p(s))
    return st


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


@pytest.mark.parametrize("prec", [_lib.PF_C128, _lib.PF_C64])
def test_noalloc_fft_matches_pf_fft2d_on_two_calls_in_a_row(prec):
    dt = torch.complex128 if prec == _lib.PF_C128 else torch.complex64
    tol = 1e-12 if prec == _lib.PF_C128 else 1e-5
    g = torch.Generator().manual_seed(5)
    # the second call is smaller than the first: the shared array holds the first call's data
    for (Nv, Nu), sign in (((1024, 2048) if prec == _lib.PF_C64 else (1024, 1024), -1), ((64, 128), +1)):
        assert Nv * Nu * (16 if prec == _lib.PF_C128 else 8) <= LIMIT
        x = torch.randn(Nv * Nu, dtype=dt, generator=g).cuda()
        ref = torch.empty_like(x)
        out = torch.empty_like(x)
        assert fft("pf_fft2d", Nv, Nu, prec, sign, x, ref) == 0, _lib.lib().pf_last_error()
        assert fft("pf_fft2d_noalloc", Nv, Nu, prec, sign, x, out) == 0, _lib.lib().pf_last_error()
        torch.cuda.synchronize()
        assert rel(out, ref) < tol, (Nv, Nu, sign, rel(out, ref))
        # and against torch's FFT in FP64 on the CPU (unitary: norm="ortho"; sign -1 is the forward)
        xc = x.cpu().to(torch.complex128).reshape(Nv, Nu)
        t = torch.fft.fft2(xc, norm="ortho") if sign < 0 else torch.fft.ifft2(xc, norm="ortho")
        assert rel(out.cpu().to(torch.complex128), t.reshape(-1)) < tol


def test_noalloc_fft_in_place_and_round_trip():
    Nv, Nu = 256, 512
    x = torch.randn(Nv * Nu, dtype=torch.complex128, generator=torch.Generator().manual_seed(9)).cuda()
    y = x.clone()
    assert fft("pf_fft2d_noalloc", Nv, Nu, _lib.PF_C128, -1, y, y) == 0
    assert fft("pf_fft2d_noalloc", Nv, Nu, _lib.PF_C128, +1, y, y) == 0
    torch.cuda.synchronize()
    assert rel(y, x) < 1e-13


def test_noalloc_fft_rejects_sizes_beyond_its_limit():
    x = torch.zeros(16, dtype=torch.complex128, device="cuda")
    L = _lib.lib()
    # 2048 x 2048 complex128 is 64 MiB: over the limit, refused before any launch
    assert fft("pf_fft2d_noalloc", 2048, 2048, _lib.PF_C128, -1, x, x) == _lib.PF_ERR_ARG
    assert b"allocation-free limit" in L.pf_last_error()
    assert fft("pf_fft2d_noalloc", 48, 64, _lib.PF_C64, -1, x, x) == _lib.PF_ERR_UNSUPPORTED
    assert fft("pf_fft2d_noalloc", 64, 64, 7, -1, x, x) == _lib.PF_ERR_ARG
