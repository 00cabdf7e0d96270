"""purify_amd -- MI355X-native PURIFY measurement operator (Phi = W G F Z S) behind PURIFY's interface.

The compute runs in libpurify_amd.so (hand-written HIP for gfx950) through the C-ABI of
include/purify_amd.h; this package is the host-side mirror of the reference's C++ interface.
"""
from . import _lib
from .kernels import kernel, kernel_from_string
from .operators import (GridDegridOperator, LinearTransform, MeasurementOperator, PsfOperator,
                        distributed_measurement_operator, measurement_operator_factory, measurementoperator,
                        operators)
from .utilities import vis_params, vis_units
from . import algorithms, pfitsio, utilities, wavelets, wproj_utilities

__all__ = ["kernel", "kernel_from_string", "LinearTransform", "MeasurementOperator", "measurementoperator",
           "operators", "GridDegridOperator", "PsfOperator",
           "measurement_operator_factory", "distributed_measurement_operator", "vis_params", "vis_units",
           "utilities", "pfitsio", "fft2d", "kernel_eval", "version", "cache_stats", "empty_cache"]


def version():
    return _lib.lib().pf_version().decode()


def cache_stats():
    """(bytes held by live handles, bytes cached for reuse) in the library's device block cache."""
    import ctypes as C
    live, cached = C.c_int64(), C.c_int64()
    _lib.check(_lib.lib().pf_cache_stats(C.byref(live), C.byref(cached)))
    return live.value, cached.value


def empty_cache():
    """Release the library's cached device blocks (like torch.cuda.empty_cache for torch's); returns the bytes
    released."""
    import ctypes as C
    n = C.c_int64()
    _lib.check(_lib.lib().pf_cache_release(C.byref(n)))
    return n.value


def fft2d(x, Nv, Nu, sign=-1):
    """Unitary 2-D FFT of a torch CUDA tensor (complex64/complex128) of Nv*Nu entries (operators.h:209-261)."""
    import ctypes as C
    import torch
    if x.dtype not in (torch.complex64, torch.complex128):
        raise TypeError(f"fft2d takes complex64 or complex128, got {x.dtype}")
    prec = _lib.PF_C128 if x.dtype == torch.complex128 else _lib.PF_C64
    from .operators import check_device_tensor
    x = check_device_tensor(x.contiguous(), int(Nv) * int(Nu), x.dtype, None, "input")
    out = torch.empty_like(x)
    s = torch.cuda.current_stream(x.device).cuda_stream
    _lib.check(_lib.lib().pf_fft2d(Nv, Nu, prec, sign, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), None,
                                   C.c_void_p(s)))
    return out


def kernel_eval(kind, x, J, which=0, os_=2.0, ftsize=1, precision="c128"):
    """Evaluate a gridding kernel on the GPU (which=0) or its correction FT (which=1, host)."""
    import numpy as np
    x = np.ascontiguousarray(x, np.float64).ravel()
    out = np.empty_like(x)
    prec = _lib.PF_C128 if precision in ("c128", "complex128") else _lib.PF_C64
    _lib.check(_lib.lib().pf_kernel_eval(int(kernel.coerce(kind)), int(which), int(J), float(os_), int(ftsize), prec,
                                         x.ctypes.data, out.ctypes.data, x.size))
    return out
