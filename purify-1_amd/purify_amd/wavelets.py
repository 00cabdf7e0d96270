"""The sparsity dictionary Psi on the GPU: factory::wavelet_operator_factory
(cpp/purify/wavelet_operator_factory.h:20-45), i.e. sopt::wavelets::SARA(wavelets) wrapped by
sopt::linear_transform(sara, imsizey, imsizex).

``wavelets`` is a list of (name, levels) with names "Dirac" or "DB1".."DB10", as the reference spells
them (tests/wavelet_factory.cc:11-14).  The LinearTransform's direct operator is the synthesis Psi
(len(wavelets) stacked coefficient images -> image) and its adjoint the analysis Psi^dagger.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .operators import check_device_tensor


def parse_wavelet(name):
    n = name.strip()
    if n.lower() == "dirac":
        return 0
    if n.upper().startswith("DB") and n[2:].isdigit():
        k = int(n[2:])
        if 1 <= k <= _lib.PF_DB_MAX:
            return k
    raise ValueError(f"unknown wavelet {name!r}: use 'Dirac' or 'DB1'..'DB{_lib.PF_DB_MAX}'")


def daubechies_filter(order):
    """The library's Daubechies scaling filter (2 * order taps, sum sqrt(2))."""
    buf = np.zeros(2 * _lib.PF_DB_MAX, np.float64)
    n = C.c_int32()
    _lib.check(_lib.lib().pf_wavelet_filter(int(order), buf.ctypes.data, C.byref(n)))
    return buf[: n.value].copy()


def _torch():
    import torch
    return torch


class WaveletOperator:
    def __init__(self, wavelets, imsizey, imsizex, precision="c64"):
        self.entries = [(parse_wavelet(nm), int(lv)) for nm, lv in wavelets]
        self.imsizey, self.imsizex = int(imsizey), int(imsizex)
        self.precision = _lib.PF_C128 if precision in ("c128", "complex128", _lib.PF_C128) else _lib.PF_C64
        db = np.array([e[0] for e in self.entries], np.int32)
        lv = np.array([e[1] for e in self.entries], np.int32)
        self._h = C.c_void_p()
        self._streams = _lib.StreamLog()
        _lib.check(_lib.lib().pf_wavelet_create(self.imsizey, self.imsizex, len(self.entries), db.ctypes.data,
                                                lv.ctypes.data, self.precision, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def size(self):
        """Number of dictionary entries (sopt::wavelets::SARA::size)."""
        return len(self.entries)

    def close(self):
        if self._h:
            self._streams.sync()
            _lib.lib().pf_wavelet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _dtype(self):
        torch = _torch()
        return torch.complex64 if self.precision == _lib.PF_C64 else torch.complex128

    def _run(self, fn, x, nin, nout, out=None, stream=None):
        if hasattr(x, "is_cuda") and x.is_cuda:
            torch = _torch()
            if x.dtype != self._dtype():
                raise TypeError(f"device input must be {self._dtype()}")
            x = x.contiguous()
            if x.numel() != nin:
                raise ValueError(f"input has {x.numel()} entries, expected {nin}")
            if out is None:
                out = torch.empty(nout, dtype=x.dtype, device=x.device)
            else:
                check_device_tensor(out, nout, x.dtype, x.device, "output")
            s = stream if stream is not None else torch.cuda.current_stream(x.device).cuda_stream
            self._streams.add(s, x.device)
            _lib.check(fn(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), _lib.PF_DEVICE,
                          _lib.PF_DEVICE, C.c_void_p(s)))
            return out
        xh = np.ascontiguousarray(np.asarray(x), np.complex128).ravel()
        if xh.size != nin:
            raise ValueError(f"input has {xh.size} entries, expected {nin}")
        yh = np.empty(nout, np.complex128)
        _lib.check(fn(self._h, xh.ctypes.data, yh.ctypes.data, _lib.PF_HOST, _lib.PF_HOST, None))
        return yh

    def synthesis(self, coeffs, out=None, stream=None):
        n = self.imsizey * self.imsizex
        return self._run(_lib.lib().pf_wavelet_forward, coeffs, n * self.size(), n, out, stream)

    def analysis(self, image, out=None, stream=None):
        n = self.imsizey * self.imsizex
        return self._run(_lib.lib().pf_wavelet_adjoint, image, n, n * self.size(), out, stream)

    def __mul__(self, coeffs):
        return self.synthesis(coeffs)

    def adjoint(self):
        return _WaveletAdjoint(self)


class _WaveletAdjoint:
    def __init__(self, w):
        self.w = w

    def __mul__(self, image):
        return self.w.analysis(image)

    def adjoint(self):
        return self.w


class distributed_wavelet_operator:  # noqa: N801 -- wavelet_operator_factory.h:19
    serial = "serial"
    mpi_sara = "mpi_sara"


def wavelet_operator_factory(distribute, wavelets, imsizey, imsizex, precision="c64"):
    """factory::wavelet_operator_factory (wavelet_operator_factory.h:20-45); the MPI-distributed SARA
    has no single-GPU meaning and raises as the reference does without MPI."""
    if distribute == distributed_wavelet_operator.serial:
        return WaveletOperator(wavelets, imsizey, imsizex, precision)
    raise RuntimeError("Distributed method not found for Wavelet Operator. Are you sure you compiled with MPI?")
