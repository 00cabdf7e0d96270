"""ctypes binding of the C-ABI in include/purify_amd.h (libpurify_amd.so, built for gfx950).

The library is required: importing this module without the built shared object raises, so no code
path can silently fall back to a CPU implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpurify_amd.so")

PF_OK, PF_ERR_ARG, PF_ERR_HIP, PF_ERR_UNSUPPORTED, PF_ERR_NOMEM, PF_ERR_STATE = 0, -1, -2, -3, -4, -5
PF_C64, PF_C128 = 0, 1
PF_HOST, PF_DEVICE = 0, 1
PF_DB_MAX = 10  # purify_amd.h: Daubechies orders DB1..DB10
PF_PADMM_IMAGE, PF_PADMM_RESIDUAL = 0, 1


class pf_padmm_params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("epsilon", C.c_double), ("nu", C.c_double),
                ("lagrange_update_scale", C.c_double), ("relative_variation", C.c_double),
                ("residual_tolerance", C.c_double), ("itermax", C.c_int32),
                ("tight_frame", C.c_int32), ("l1_itermax", C.c_int32), ("l1_tolerance", C.c_double),
                ("l1_nu", C.c_double), ("positivity", C.c_int32), ("real_constraint", C.c_int32)]


class pf_fb_params(C.Structure):
    _fields_ = [("mu", C.c_double), ("beta", C.c_double), ("sigma", C.c_double), ("relative_variation", C.c_double),
                ("itermax", C.c_int32), ("tight_frame", C.c_int32), ("l1_itermax", C.c_int32),
                ("l1_tolerance", C.c_double), ("l1_nu", C.c_double), ("positivity", C.c_int32),
                ("real_constraint", C.c_int32)]


class pf_padmm_result(C.Structure):
    _fields_ = [("niters", C.c_int32), ("converged", C.c_int32), ("relative_variation", C.c_double),
                ("residual_norm", C.c_double), ("gamma", C.c_double), ("l1_iters_last", C.c_int32),
                ("l1_iters_total", C.c_int64)]
PF_LAMBDA, PF_RADIANS, PF_PIXELS = 0, 1, 2


class pf_op_desc(C.Structure):
    _fields_ = [
        ("imsizey", C.c_int32), ("imsizex", C.c_int32),
        ("oversample_ratio", C.c_double),
        ("kernel", C.c_int32),
        ("Ju", C.c_int32), ("Jv", C.c_int32),
        ("w_term", C.c_int32), ("Jw", C.c_int32),
        ("cellx", C.c_double), ("celly", C.c_double),
        ("units", C.c_int32), ("precision", C.c_int32),
        ("M", C.c_int64),
        ("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p),
        ("weights", C.c_void_p),
        ("power_iters", C.c_int32), ("power_tol", C.c_double),
        ("tile_u", C.c_int32), ("tile_v", C.c_int32),
        ("energy_chirp", C.c_double), ("energy_wproj", C.c_double),
    ]


class pf_op_info(C.Structure):
    _fields_ = [
        ("imsizey", C.c_int32), ("imsizex", C.c_int32), ("Nv", C.c_int32), ("Nu", C.c_int32),
        ("Ju", C.c_int32), ("Jv", C.c_int32), ("precision", C.c_int32), ("w_term", C.c_int32),
        ("M", C.c_int64), ("nnz_w", C.c_int64),
        ("tile_u", C.c_int32), ("tile_v", C.c_int32), ("n_tiles", C.c_int32),
        ("tile_list_len", C.c_int64),
        ("norm", C.c_double),
        ("device_bytes", C.c_int64),
    ]


class pf_csr(C.Structure):
    _fields_ = [
        ("rows", C.c_int64), ("cols", C.c_int64), ("nnz", C.c_int64),
        ("rowptr", C.POINTER(C.c_int64)), ("colidx", C.POINTER(C.c_int32)), ("vals", C.POINTER(C.c_double)),
    ]


_lib = None

# name -> (restype, argtypes); kept in the order of include/purify_amd.h
SIGNATURES = {
    "pf_op_desc_init": (None, [C.POINTER(pf_op_desc)]),
    "pf_op_create": (C.c_int, [C.POINTER(pf_op_desc), C.POINTER(C.c_void_p)]),
    "pf_op_destroy": (C.c_int, [C.c_void_p]),
    "pf_op_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "pf_op_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "pf_op_apply_normal": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "pf_padmm_params_init": (C.c_int, [C.c_void_p]),
    "pf_padmm_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "pf_padmm_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "pf_padmm_reset": (C.c_int, [C.c_void_p]),
    "pf_padmm_enqueue": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "pf_fb_params_init": (C.c_int, [C.c_void_p]),
    "pf_super_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "pf_super_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "pf_super_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_super_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_super_destroy": (C.c_int, [C.c_void_p]),
    "pf_fb_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "pf_padmm_sync": (C.c_int, [C.c_void_p, C.c_void_p]),
    "pf_padmm_get": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int]),
    "pf_padmm_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "pf_padmm_destroy": (C.c_int, [C.c_void_p]),
    "pf_wavelet_filter": (C.c_int, [C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]),
    "pf_wavelet_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                    C.POINTER(C.c_void_p)]),
    "pf_wavelet_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "pf_wavelet_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "pf_wavelet_destroy": (C.c_int, [C.c_void_p]),
    "pf_psf_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_void_p, C.POINTER(C.c_void_p)]),
    "pf_psf_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "pf_psf_norm": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pf_psf_destroy": (C.c_int, [C.c_void_p]),
    "pf_op_power_method": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_void_p, C.POINTER(C.c_double)]),
    "pf_op_set_norm": (C.c_int, [C.c_void_p, C.c_double]),
    "pf_op_get_norm": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "pf_op_get_info": (C.c_int, [C.c_void_p, C.POINTER(pf_op_info)]),
    "pf_op_apply_FZ": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_op_apply_FZ_adjoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_op_apply_G": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_op_apply_GH": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_fft2d": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p]),
    "pf_fft2d_noalloc": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "pf_kernel_eval": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, C.c_int32, C.c_void_p,
                                 C.c_void_p, C.c_int64]),
    "pf_op_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "pf_op_get_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]),
    "pf_op_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]),
    "pf_op_profile_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "pf_op_profile_get": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_double),
                                    C.POINTER(C.c_int64)]),
    "pf_synth_uv": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_void_p, C.c_void_p]),
    "pf_synth_image": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "pf_synth_cnormal": (C.c_int, [C.c_uint64, C.c_int64, C.c_double, C.c_void_p]),
    "pf_wprojection_matrix": (C.c_int, [C.POINTER(pf_csr), C.c_int32, C.c_int32, C.c_void_p, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.POINTER(pf_csr)]),
    "pf_csr_free": (None, [C.POINTER(pf_csr)]),
    "pf_op_create_csr": (C.c_int, [C.POINTER(pf_op_desc), C.POINTER(pf_csr), C.POINTER(C.c_void_p)]),
    "pf_padmm_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]),
    "pf_padmm_batch_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "pf_padmm_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "pf_padmm_batch_destroy": (C.c_int, [C.c_void_p]),
    "pf_cache_stats": (C.c_int, [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pf_cache_release": (C.c_int, [C.POINTER(C.c_int64)]),
    "pf_last_error": (C.c_char_p, []),
    "pf_version": (C.c_char_p, []),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"purify_amd HIP library not built ({LIB_PATH}); run python purify-1_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class PurifyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check(rc):
    if rc != PF_OK:
        msg = lib().pf_last_error().decode(errors="replace")
        raise PurifyError(rc, msg)
    return rc


class StreamLog:
    """The streams a handle's device calls were queued on.  Destroying a handle returns its device memory to
    the library's block cache without waiting for the device, so the handle's close() first synchronises
    every stream logged here (each one only: other streams' work goes on)."""

    def __init__(self):
        self._s = set()

    def add(self, stream, device):
        self._s.add((int(stream or 0), 0 if device.index is None else device.index))

    def sync(self):
        if not self._s:
            return
        import torch
        for s, d in self._s:
            if s == 0:
                torch.cuda.default_stream(d).synchronize()
            else:
                torch.cuda.ExternalStream(s, device=torch.device("cuda", d)).synchronize()
        self._s.clear()
