"""FP64 CPU oracle for PURIFY's measurement operator -- TEST INFRASTRUCTURE ONLY.

This package wraps ``oracle/liboracle.so`` (built from ``purify_oracle.c`` by ``oracle/Makefile``)
with ctypes and numpy.  It is the checker: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it.  The product library under ``purify-1_amd/`` never
imports, links or calls it.

Every function restates a reference function; the C file cites reference file:line for each.
Complex arrays are numpy complex128 (interleaved re/im doubles, as the C side expects).

Parity pins (what each function is checked against) live in ``tests/test_oracle_*.py``.  The SOPT
power-method iteration is restated from the published power-iteration algorithm; its stopping
rule is "parity unpinned" (SOPT is not in the reference tree) and only its converged value is
pinned, against the largest singular value of the dense operator.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

KB, GAUSS, BOX, PSWF, KBMIN, GAUSS_ALT = range(6)
KERNELS = {"kb": KB, "gauss": GAUSS, "box": BOX, "pswf": PSWF, "kbmin": KBMIN, "gauss_alt": GAUSS_ALT}
W_APPROX, W_BOX, W_SPHERE = range(3)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

_lib = None


def build() -> None:
    """Compile liboracle.so with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    d, i, i64, vp = C.c_double, C.c_int, C.c_int64, C.c_void_p
    sig = {
        "orc_bessel_i0": (d, [d]),
        "orc_kaiser_bessel": (d, [d, i]),
        "orc_ft_kaiser_bessel": (d, [d, i]),
        "orc_gaussian": (d, [d, i]),
        "orc_ft_gaussian": (d, [d, i]),
        "orc_pswf": (d, [d, i]),
        "orc_ft_pswf": (d, [d, i]),
        "orc_kernel_eval": (d, [i, i, d, i, i, d]),
        "orc_correction2d": (i, [d, i, i, i, i, i, _dp]),
        "orc_zero_pad": (None, [_dp, i, i, d, _cp, _cp]),
        "orc_zero_pad_adjoint": (None, [_dp, i, i, d, _cp, _cp]),
        "orc_fft2d": (None, [i, i, _cp, _cp, i]),
        "orc_fft2d_raw": (None, [i, i, _cp, _cp, i]),
        "orc_gridding_matrix_nnz": (i64, [i64, i, i, d, i, i, i, i]),
        "orc_gridding_matrix": (i64, [_dp, _dp, _cp, i64, i, i, d, i, i, i, i, i, i, _i64p, _i32p, _cp]),
        "orc_gridding_matrix_w": (i64, [_dp, _dp, _dp, _cp, i64, i, i, d, i, i, i, i, i, i, i, d, d,
                                        _i64p, _i32p, _cp]),
        "orc_projection": (None, [i, i, i, i, i, i, d, d, i, i, d, d, d, d, _cp]),
        "orc_csr_adjoint": (None, [i64, i64, _i64p, _i32p, _cp, _i64p, _i32p, _cp]),
        "orc_spmv": (None, [i64, _i64p, _i32p, _cp, _cp, _cp]),
        "orc_wkernel_eval": (None, [i, d, d, d, d, d, i, i, d, _dp]),
        "orc_dwt2": (None, [_cp, i, i, _dp, i, i, _cp]),
        "orc_idwt2": (None, [_cp, i, i, _dp, i, i, _cp]),
        "orc_pixel_to_lambda": (d, [d, i, d]),
        "orc_linear_convol_1d": (None, [_cp, i, _cp, i, _cp]),
        "orc_linear_convol_2d": (None, [_cp, i, _cp, i, _cp, i, i, _cp]),
        "orc_generate_chirp": (None, [d, d, d, i, i, _cp]),
        "orc_sparsify_row_dense_thres": (d, [_cp, i64, d]),
        "orc_sparsify_row_thres_real": (d, [_dp, i64, d]),
        "orc_row_wise_convolution": (None, [_i32p, _cp, i, _cp, i, i, _cp]),
        "orc_wprojection_matrix_dense": (i, [i64, _i64p, _i32p, _cp, i, i, _dp, d, d, d, d, _cp, _u8p]),
        "orc_uv_to_pixels": (None, [_dp, _dp, i64, i, d, d, i, i, d, _dp, _dp]),
        "orc_calculate_l2_radius": (d, [i64, d, d, i]),
        "orc_snr_to_standard_deviation": (d, [_cp, i64, d]),
        "orc_op_create": (vp, [_dp, _dp, _dp, _cp, i64, i, i, d, i, i, i, i, i, d, d, i]),
        "orc_op_destroy": (None, [vp]),
        "orc_op_set_norm": (None, [vp, d]),
        "orc_op_get_norm": (d, [vp]),
        "orc_op_nnz": (i64, [vp]),
        "orc_op_forward": (None, [vp, _cp, _cp]),
        "orc_op_adjoint": (None, [vp, _cp, _cp]),
        "orc_op_power_method": (d, [vp, i, d, vp]),
        "orc_op_apply_G": (None, [vp, _cp, _cp]),
        "orc_op_apply_GH": (None, [vp, _cp, _cp]),
        "orc_omp_threads": (i, []),
        "orc_set_omp_threads": (None, [i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- scalar kernels
def kernel_eval(kind, which, x, J, ftsize=1, os_=2.0):
    L = lib()
    return np.array([L.orc_kernel_eval(int(kind), int(which), float(t), int(J), int(ftsize), float(os_))
                     for t in np.ravel(x)])


def kaiser_bessel(x, J):
    return np.array([lib().orc_kaiser_bessel(float(t), int(J)) for t in np.ravel(x)])


def ft_kaiser_bessel(x, J):
    return np.array([lib().orc_ft_kaiser_bessel(float(t), int(J)) for t in np.ravel(x)])


def gaussian(x, J):
    return np.array([lib().orc_gaussian(float(t), int(J)) for t in np.ravel(x)])


def ft_gaussian(x, J):
    return np.array([lib().orc_ft_gaussian(float(t), int(J)) for t in np.ravel(x)])


def pswf(x, J):
    return np.array([lib().orc_pswf(float(t), int(J)) for t in np.ravel(x)])


def ft_pswf(x, J):
    return np.array([lib().orc_ft_pswf(float(t), int(J)) for t in np.ravel(x)])


def bessel_i0(x):
    return np.array([lib().orc_bessel_i0(float(t)) for t in np.ravel(x)])


# ---------------------------------------------------------------- operator pieces
def grid_dims(ny, nx, os_):
    return int(np.floor(ny * os_)), int(np.floor(nx * os_))


def correction2d(os_, ny, nx, kind=KB, Ju=4, Jv=4):
    S = np.zeros(ny * nx)
    if lib().orc_correction2d(float(os_), ny, nx, kind, Ju, Jv, S) != 0:
        raise RuntimeError("Incorrect input: PSWF requires a support of 6")
    return S


def zero_pad(S, ny, nx, os_, x):
    Nv, Nu = grid_dims(ny, nx, os_)
    g = np.zeros(Nv * Nu, np.complex128)
    lib().orc_zero_pad(_d(S), ny, nx, float(os_), _c(x), g)
    return g


def zero_pad_adjoint(S, ny, nx, os_, g):
    x = np.zeros(ny * nx, np.complex128)
    lib().orc_zero_pad_adjoint(_d(S), ny, nx, float(os_), _c(g), x)
    return x


def fft2d(Nv, Nu, x, sign=-1, unitary=True):
    out = np.zeros(Nv * Nu, np.complex128)
    (lib().orc_fft2d if unitary else lib().orc_fft2d_raw)(Nv, Nu, _c(x), out, int(sign))
    return out


class CSR:
    def __init__(self, rows, cols, rowptr, colidx, vals):
        self.rows, self.cols = rows, cols
        self.rowptr, self.colidx, self.vals = rowptr, colidx, vals

    @property
    def nnz(self):
        return int(self.rowptr[-1])

    def dot(self, x):
        y = np.zeros(self.rows, np.complex128)
        lib().orc_spmv(self.rows, self.rowptr, self.colidx, self.vals, _c(x), y)
        return y

    def adjoint(self):
        tr = np.zeros(self.cols + 1, np.int64)
        tc = np.zeros(max(self.nnz, 1), np.int32)
        tv = np.zeros(max(self.nnz, 1), np.complex128)
        lib().orc_csr_adjoint(self.rows, self.cols, self.rowptr, self.colidx, self.vals, tr, tc, tv)
        return CSR(self.cols, self.rows, tr, tc[: self.nnz], tv[: self.nnz])

    def todense(self):
        D = np.zeros((self.rows, self.cols), np.complex128)
        for r in range(self.rows):
            for k in range(self.rowptr[r], self.rowptr[r + 1]):
                D[r, self.colidx[k]] += self.vals[k]
        return D


def gridding_matrix(u, v, weights, ny, nx, os_, kind=KB, Jeval_u=4, Jeval_v=4, Ju=4, Jv=4, literal_phase=False):
    M = len(u)
    nnz = lib().orc_gridding_matrix_nnz(M, ny, nx, float(os_), Ju, Jv, 0, 0)
    rp = np.zeros(M + 1, np.int64)
    ci = np.zeros(max(nnz, 1), np.int32)
    vv = np.zeros(max(nnz, 1), np.complex128)
    lib().orc_gridding_matrix(_d(u), _d(v), _c(weights), M, ny, nx, float(os_), kind, Jeval_u, Jeval_v, Ju, Jv,
                              int(literal_phase), rp, ci, vv)
    Nv, Nu = grid_dims(ny, nx, os_)
    return CSR(M, Nv * Nu, rp, ci[:nnz], vv[:nnz])


def gridding_matrix_w(u, v, w, weights, ny, nx, os_, kind=KB, Jeval_u=4, Jeval_v=4, Ju=4, Jv=4, Jw=6,
                      wkind=W_APPROX, cellx=1.0, celly=1.0):
    M = len(u)
    nnz = lib().orc_gridding_matrix_nnz(M, ny, nx, float(os_), Ju, Jv, 1, Jw)
    rp = np.zeros(M + 1, np.int64)
    ci = np.zeros(max(nnz, 1), np.int32)
    vv = np.zeros(max(nnz, 1), np.complex128)
    lib().orc_gridding_matrix_w(_d(u), _d(v), _d(w), _c(weights), M, ny, nx, float(os_), kind, Jeval_u, Jeval_v,
                                Ju, Jv, Jw, wkind, float(cellx), float(celly), rp, ci, vv)
    Nv, Nu = grid_dims(ny, nx, os_)
    return CSR(M, Nv * Nu, rp, ci[:nnz], vv[:nnz])


def projection(u, v, w, kind=KB, Jeval_u=4, Ju=4, Jv=4, Jw=6, wkind=W_APPROX, cellx=1.0, celly=1.0, nx=16, ny=16,
               os_=2.0):
    out = np.zeros((Jv + Jw - 1) * (Ju + Jw - 1), np.complex128)
    lib().orc_projection(kind, Jeval_u, Ju, Jv, Jw, wkind, float(cellx), float(celly), nx, ny, float(os_),
                         float(u), float(v), float(w), out)
    return out.reshape(Jv + Jw - 1, Ju + Jw - 1)


def wkernel(u, v, w, cellx=1.0, celly=1.0, nx=16, ny=16, os_=2.0, wkind=None):
    """One w-kernel sample at pixel offsets (u, v) (projection_kernels.cc:8-58); wkind defaults to the
    approximate w-projection kernel w_projection_kernel_approx (projection_kernels.cc:31-47)."""
    out = np.zeros(2)
    lib().orc_wkernel_eval(W_APPROX if wkind is None else wkind, float(u), float(v), float(w), float(cellx),
                           float(celly), int(nx), int(ny), float(os_), out)
    return complex(out[0], out[1])


def pixel_to_lambda(cell, imsize, os_=2.0):
    """projection_kernels.cc:28-30: wavelengths per grid pixel."""
    return float(lib().orc_pixel_to_lambda(float(cell), int(imsize), float(os_)))


def linear_convol_1d(f, g_padded):
    f, g = _c(f), _c(g_padded)
    out = np.zeros(len(g) - len(f) + 1, np.complex128)
    lib().orc_linear_convol_1d(f, len(f), g, len(g), out)
    return out


def linear_convol_2d(fu, fv, g):
    fu, fv, g = _c(fu), _c(fv), _c(g)
    Jgv, Jgu = g.shape
    out = np.zeros((len(fv) + Jgv - 1, len(fu) + Jgu - 1), np.complex128)
    lib().orc_linear_convol_2d(fu, len(fu), fv, len(fv), np.ascontiguousarray(g).ravel(), Jgv, Jgu, out.ravel())
    return out


def generate_chirp(w_rate, cell_x, cell_y, x_size, y_size):
    out = np.zeros(x_size * y_size, np.complex128)
    lib().orc_generate_chirp(float(w_rate), float(cell_x), float(cell_y), x_size, y_size, out)
    return out


def sparsify_row_dense_thres(row, energy):
    r = _c(row)
    return lib().orc_sparsify_row_dense_thres(r, len(r), float(energy))


def sparsify_row_thres(absvals, energy):
    a = _d(absvals)
    return lib().orc_sparsify_row_thres_real(a, len(a), float(energy))


def row_wise_convolution(g_idx, g_val, chirp_dense, x_size, y_size):
    out = np.zeros(x_size * y_size, np.complex128)
    lib().orc_row_wise_convolution(np.ascontiguousarray(g_idx, np.int32), _c(g_val), len(g_idx), _c(chirp_dense),
                                   x_size, y_size, out)
    return out


def wprojection_matrix_dense(G: CSR, x_size, y_size, w, cell_x, cell_y, e_chirp, e_wproj):
    N = x_size * y_size
    out = np.zeros(G.rows * N, np.complex128)
    mask = np.zeros(G.rows * N, np.uint8)
    lib().orc_wprojection_matrix_dense(G.rows, G.rowptr, G.colidx, _c(G.vals), x_size, y_size, _d(w),
                                       float(cell_x), float(cell_y), float(e_chirp), float(e_wproj), out, mask)
    return out.reshape(G.rows, N), mask.reshape(G.rows, N).astype(bool)


def uv_to_pixels(u, v, units, cell_x, cell_y, nx, ny, os_):
    u, v = _d(u), _d(v)
    up, vp = np.zeros_like(u), np.zeros_like(v)
    lib().orc_uv_to_pixels(u, v, len(u), int(units), float(cell_x), float(cell_y), nx, ny, float(os_), up, vp)
    return up, vp


def calculate_l2_radius(y_size, sigma=0.0, n_sigma=2.0, distribution="chi"):
    return lib().orc_calculate_l2_radius(int(y_size), float(sigma), float(n_sigma), int(distribution == "chi^2"))


def snr_to_standard_deviation(y0, snr=30.0):
    y = _c(y0)
    return lib().orc_snr_to_standard_deviation(y, len(y), float(snr))


class Operator:
    """Phi = G F Z S / norm with the explicit CSR G and G^H (operators.h:298-414)."""

    def __init__(self, u_px, v_px, w, weights, ny, nx, os_=2.0, kernel=KB, Ju=4, Jv=4, w_term=False, Jw=6,
                 cellx=1.0, celly=1.0, literal_phase=False, power_iters=0, power_tol=1e-4):
        self.ny, self.nx, self.os = ny, nx, float(os_)
        self.M = len(u_px)
        w = np.zeros(self.M) if w is None else w
        weights = np.ones(self.M, np.complex128) if weights is None else weights
        self._h = lib().orc_op_create(_d(u_px), _d(v_px), _d(w), _c(weights), self.M, ny, nx, self.os, int(kernel),
                                      Ju, Jv, int(w_term), Jw, float(cellx), float(celly), int(literal_phase))
        if not self._h:
            raise RuntimeError("Incorrect input: PSWF requires a support of 6")
        self.norm = 1.0
        if power_iters > 0:
            self.norm = self.power_method(power_iters, power_tol)
            lib().orc_op_set_norm(self._h, self.norm)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_op_destroy(h)
            self._h = None

    @property
    def nnz(self):
        return lib().orc_op_nnz(self._h)

    def set_norm(self, norm):
        self.norm = float(norm)
        lib().orc_op_set_norm(self._h, self.norm)

    def forward(self, x):
        y = np.zeros(self.M, np.complex128)
        lib().orc_op_forward(self._h, _c(x), y)
        return y

    def adjoint(self, y):
        x = np.zeros(self.nx * self.ny, np.complex128)
        lib().orc_op_adjoint(self._h, _c(y), x)
        return x

    def apply_G(self, grid):
        y = np.zeros(self.M, np.complex128)
        lib().orc_op_apply_G(self._h, _c(grid), y)
        return y

    def apply_GH(self, y):
        Nv, Nu = grid_dims(self.ny, self.nx, self.os)
        g = np.zeros(Nv * Nu, np.complex128)
        lib().orc_op_apply_GH(self._h, _c(y), g)
        return g

    def power_method(self, iters, tol, start=None, normalised=False):
        """||Phi||_2 of the raw operator, or of Phi/norm when normalised=True."""
        old = self.norm
        if not normalised:
            lib().orc_op_set_norm(self._h, 1.0)
        if start is None:
            r = lib().orc_op_power_method(self._h, int(iters), float(tol), None)
        else:
            s = _c(start)
            r = lib().orc_op_power_method(self._h, int(iters), float(tol), s.ctypes.data_as(C.c_void_p))
        lib().orc_op_set_norm(self._h, old)
        return r


def power_method_fn(apply, start, iters, tol):
    """SOPT power_method on a LinearTransform {A, identity} (as the compact operators call it,
    compact_operators.h:164-165 and 298-299): x <- A x / ||A x||, stopping when
    |lambda_k - lambda_(k-1)| / lambda_(k-1) < tol^2; returns sqrt(lambda).  The same iteration as
    orc_op_power_method; its stopping rule is parity unpinned (SOPT is not in the reference tree)."""
    if iters <= 0:
        return 1.0
    x = np.asarray(start, np.complex128).copy()
    x /= np.linalg.norm(x)
    old = 0.0
    for _ in range(iters):
        x = apply(x)
        val = float(np.linalg.norm(x))
        if not val > 0:
            old = val
            break
        x /= val
        if old > 0 and tol * tol > abs(old - val) / old:
            old = val
            break
        old = val
    return float(np.sqrt(old))


def gtg_matrix(G: CSR, prune=1e-7):
    """G^H G with entries of |value| <= prune removed (compact_operators.h:104-113, Eigen's
    prune with the reference's 1e-7 threshold), as a scipy CSR matrix."""
    import scipy.sparse as sp
    Gs = sp.csr_matrix((G.vals, G.colidx, G.rowptr), shape=(G.rows, G.cols))
    GTG = (Gs.conj().T @ Gs).tocsr()
    if prune is not None and prune > 0:
        GTG.data[np.abs(GTG.data) <= prune] = 0
        GTG.eliminate_zeros()
    return GTG


def grid_degrid(u_px, v_px, weights, ny, nx, os_, x, kind=KB, Ju=4, Jv=4, norm=1.0, prune=1e-7):
    """init_grid_degrid_operator_2d's operator (compact_operators.h:118-171) on x:
    Z^dagger F^dagger (G^H G) F Z S x / norm^2 with the explicit, pruned G^H G; G built with the
    kernel-slot swap of compact_operators.h:141-142 (the u-axis evaluates the Jv kernel)."""
    M = len(u_px)
    weights = np.ones(M, np.complex128) if weights is None else np.asarray(weights, np.complex128)
    G = gridding_matrix(u_px, v_px, weights, ny, nx, os_, kind, Jv, Ju, Ju, Jv)
    GTG = gtg_matrix(G, prune)
    S = correction2d(os_, ny, nx, kind, Ju, Jv)
    Nv, Nu = grid_dims(ny, nx, os_)
    g = fft2d(Nv, Nu, zero_pad(S, ny, nx, os_, x))
    g = GTG @ g
    return zero_pad_adjoint(S, ny, nx, os_, fft2d(Nv, Nu, g, sign=+1)) / (norm * norm)


def psf_convolve(op: "Operator", x, psf_norm=1.0):
    """init_psf_convolve_2d's operator (compact_operators.h:279-306) on x: psf = Phi^dagger Phi delta
    with delta at sub2ind(floor(ny/2) - 1, floor(nx/2) - 1) (utilities.cc:321-333: row * cols + col),
    then IFFT(FFT(psf) . FFT(x)) with unitary n x n FFTs, divided by psf_norm^2."""
    ny, nx = op.ny, op.nx
    delta = np.zeros(ny * nx, np.complex128)
    delta[(int(np.floor(ny * 0.5)) - 1) * nx + int(np.floor(nx * 0.5)) - 1] = 1.0
    psf = op.adjoint(op.forward(delta))
    ft = fft2d(ny, nx, psf)
    return fft2d(ny, nx, ft * fft2d(ny, nx, x), sign=+1) / (psf_norm * psf_norm)


# ---- wavelets (sopt::wavelets::SARA as used by wavelet_operator_factory.h:20-45) ----------------
# SOPT is not in the reference tree, so this restates the published construction: Daubechies'
# extremal-phase scaling filters (Ten Lectures on Wavelets, 1992, section 6.1) and the periodic
# down-convolution lo[i] = sum_k h[k] x[(2i + k) mod n], hi with g[k] = (-1)^k h[L-1-k], low-pass
# half first, rows then columns per level, the next level on the low-low block.  Parity against SOPT
# itself is unpinned; the tests pin the filters by their defining identities and closed forms
# (Haar, DB2), and the transform by orthonormality, perfect reconstruction and a Haar known answer.
def daubechies(N):
    """Scaling filter with N vanishing moments (2N taps, sum sqrt(2)), extremal phase."""
    from math import comb
    if N == 1:
        return np.array([1.0, 1.0]) / np.sqrt(2.0)
    P = [comb(N - 1 + k, k) for k in range(N)][::-1]  # highest degree first
    y = np.roots(P)
    z = []
    for yy in y:
        b = 4 * yy - 2
        s_ = np.sqrt(b * b - 4 + 0j)
        z1, z2 = (-b + s_) / 2, (-b - s_) / 2
        z.append(z1 if abs(z1) < abs(z2) else z2)
    z += [-1.0] * N
    c = np.real(np.poly(z))  # highest degree first
    return c * np.sqrt(2.0) / c.sum()


def _dwt_line(x, h):
    # lo[i] = sum_k h[k] x[(2i + k) mod n], hi with g; x[(2i + k) mod n] is x2[(i + k // 2) mod half, k % 2]
    # for the pairs x2 = x viewed as (half, 2)
    n, L = x.shape[-1], len(h)
    g = np.array([(-1) ** k * h[L - 1 - k] for k in range(L)])
    half = n // 2
    x2 = x.reshape(x.shape[:-1] + (half, 2))
    lo = np.zeros(x.shape[:-1] + (half,), np.result_type(x, h))
    hi = np.zeros_like(lo)
    for k in range(L):
        xs = np.roll(x2[..., k % 2], -(k // 2), axis=-1)
        lo += h[k] * xs
        hi += g[k] * xs
    return np.concatenate([lo, hi], axis=-1)


def _idwt_line(c, h):
    n, L = c.shape[-1], len(h)
    g = np.array([(-1) ** k * h[L - 1 - k] for k in range(L)])
    half = n // 2
    x = np.zeros(c.shape, np.result_type(c, h))  # C-contiguous, so x2 below is a view of it
    x2 = x.reshape(x.shape[:-1] + (half, 2))
    lo, hi = c[..., :half], c[..., half:]
    for k in range(L):
        # x[(2i + k) mod n] += h[k] lo[i] + g[k] hi[i], i.e. pair (i + k // 2) mod half, slot k % 2
        x2[..., k % 2] += np.roll(h[k] * lo + g[k] * hi, k // 2, axis=-1)
    return x


def dwt2_numpy(img, ny, nx, h, levels):
    """The periodic 2-D analysis in numpy (rows, then columns, of the shrinking top-left block per level);
    dwt2 runs the same in the C restatement (orc_dwt2), and the tests hold the two together."""
    c = np.asarray(img, np.complex128).reshape(ny, nx).copy()
    for l in range(levels):
        hh, ww = ny >> l, nx >> l
        blk = _dwt_line(c[:hh, :ww], h)
        c[:hh, :ww] = _dwt_line(blk.T, h).T
    return c.ravel()


def idwt2_numpy(coef, ny, nx, h, levels):
    c = np.asarray(coef, np.complex128).reshape(ny, nx).copy()
    for l in range(levels - 1, -1, -1):
        hh, ww = ny >> l, nx >> l
        blk = _idwt_line(c[:hh, :ww].T, h).T
        c[:hh, :ww] = _idwt_line(blk, h)
    return c.ravel()


def dwt2(img, ny, nx, h, levels):
    h = np.ascontiguousarray(h, np.float64)
    out = np.zeros(ny * nx, np.complex128)
    lib().orc_dwt2(_c(img), int(ny), int(nx), h, len(h), int(levels), out)
    return out


def idwt2(coef, ny, nx, h, levels):
    h = np.ascontiguousarray(h, np.float64)
    out = np.zeros(ny * nx, np.complex128)
    lib().orc_idwt2(_c(coef), int(ny), int(nx), h, len(h), int(levels), out)
    return out


def sara_analysis(img, ny, nx, entries):
    """Psi^dagger: entries = [(order, levels)], order 0 = Dirac; stacked, each / sqrt(len(entries))."""
    f = 1.0 / np.sqrt(len(entries))
    out = []
    for order, lev in entries:
        out.append(np.asarray(img, np.complex128).ravel() * f if order == 0 else dwt2(img, ny, nx, daubechies(order), lev) * f)
    return np.concatenate(out)


def sara_synthesis(coef, ny, nx, entries):
    """Psi: the sum of the entries' reconstructions, each / sqrt(len(entries))."""
    f = 1.0 / np.sqrt(len(entries))
    N = ny * nx
    x = np.zeros(N, np.complex128)
    for i, (order, lev) in enumerate(entries):
        ci = np.asarray(coef[i * N:(i + 1) * N], np.complex128)
        x += (ci if order == 0 else idwt2(ci, ny, nx, daubechies(order), lev)) * f
    return x


def soft_threshold(c, t):
    """soft(c, t) = c max(0, 1 - t / |c|) (complex magnitude shrinkage)."""
    m = np.abs(c)
    f = np.where(m > t, 1.0 - t / np.where(m > 0, m, 1.0), 0.0)
    return c * f


def project_c(x, positivity=True, real=True):
    """The constraint set of the iterative l1 proximal (algorithm_factory.h:65-66): real zeroes Im x,
    positivity clamps Re x and Im x at 0."""
    x = np.asarray(x, np.complex128)
    re, im = x.real.copy(), (np.zeros_like(x.real) if real else x.imag.copy())
    if positivity:
        re, im = np.maximum(re, 0.0), np.maximum(im, 0.0)
    return re + 1j * im


def l1_prox_constrained(x0, gamma, psit, psi, nu=1.0, itermax=50, tol=1e-2, positivity=True, real=True):
    """argmin_x 1/2 ||x - x0||^2 + gamma ||Psi^dagger x||_1 + i_C(x) by the dual forward-backward iteration
    (Combettes & Pesquet 2011, "Proximal splitting methods in signal processing"; Fadili & Starck
    2009) with step 1 / nu -- the l1 proximal padmm_factory configures for SARA
    (algorithm_factory.h:60-66: tolerance 1e-2, nu 1, 50 iterations, positivity and real constraints):
        u_0 = 0, x_0 = P_C(x0);  u_k = P_gamma(u_(k-1) + Psi^dagger x_(k-1) / nu),  x_k = P_C(x0 - Psi u_k)
    P_gamma clamps each coefficient's modulus to gamma.  Stops after itermax steps or when the objective
    changes by less than tol relative.  SOPT (not in the reference tree) is not followed line by line:
    parity with its iterates is unpinned.  Returns (x, steps)."""
    x0 = np.asarray(x0, np.complex128)
    x = project_c(x0, positivity, real)
    c = psit(x)
    obj = 0.5 * float(np.vdot(x - x0, x - x0).real) + gamma * float(np.abs(c).sum())
    u = np.zeros_like(c)
    k = 0
    while k < itermax:
        k += 1
        w = u + c / nu
        m = np.abs(w)
        u = w * np.where(m > gamma, gamma / np.where(m > 0, m, 1.0), 1.0)
        x = project_c(x0 - psi(u), positivity, real)
        c = psit(x)
        o = 0.5 * float(np.vdot(x - x0, x - x0).real) + gamma * float(np.abs(c).sum())
        done = abs(o - obj) < tol * abs(o)
        obj = o
        if done:
            break
    return x, k


def padmm(op: "Operator", entries, y, epsilon, iters, gamma=None, nu=1.0, rho=0.9, relative_variation=1e-3,
          residual_tolerance=None, stop=False, tight_frame=True, l1_itermax=50, l1_tol=1e-2, l1_nu=1.0,
          positivity=True, real=True):
    """Imaging proximal ADMM with a tight-frame l1 proximal (the iteration SOPT's ImagingProximalADMM runs
    with padmm_factory's parameters, algorithm_factory.h:36-81), in FP64:
        z = P_eps(-lambda - r);  x~ = x - Phi^dagger(r + lambda + z) / nu
        x = x~ + Psi(soft(Psi^dagger x~, gamma l1_nu) - Psi^dagger x~) / l1_nu
        r = Phi x - y;  lambda += rho (r + z)
    (nu is the ADMM step; l1_nu the l1 proximal's frame bound, SOPT's L1TightFrame nu)
    from x = Phi^dagger y / nu, r = Phi x - y, lambda = 0; gamma = max |Psi^dagger Phi^dagger y| * 1e-3 if
    not given.  tight_frame False: x = l1_prox_constrained(x~, gamma, ...) instead (SARA).  Restated from
    the published algorithm (SOPT is not in the reference tree): parity with SOPT's iterates is
    unpinned.  Returns (x, stats, gamma, niters): stats[k] = (||x_k - x_(k-1)||^2, ||x_k||^2, ||r_k||^2)."""
    ny, nx = op.ny, op.nx
    y = np.asarray(y, np.complex128)
    psit = lambda v: sara_analysis(v, ny, nx, entries)  # noqa: E731
    psi = lambda c: sara_synthesis(c, ny, nx, entries)  # noqa: E731
    if gamma is None or gamma <= 0:
        gamma = float(np.abs(psit(op.adjoint(y))).max()) * 1e-3
    res_tol = epsilon if not residual_tolerance else residual_tolerance
    x = op.adjoint(y) / nu
    r = op.forward(x) - y
    lam = np.zeros_like(y)
    stats = []
    it = 0
    for it in range(1, iters + 1):
        v = -lam - r
        nv = np.linalg.norm(v)
        z = v * (epsilon / nv if nv > epsilon else 1.0)
        xu = x - op.adjoint(r + lam + z) / nu
        if tight_frame:
            c = psit(xu)
            xn = xu + psi(soft_threshold(c, gamma * l1_nu) - c) / l1_nu
        else:
            xn, _ = l1_prox_constrained(xu, gamma, psit, psi, l1_nu, l1_itermax, l1_tol, positivity, real)
        r = op.forward(xn) - y
        lam = lam + rho * (r + z)
        stats.append((float(np.vdot(xn - x, xn - x).real), float(np.vdot(xn, xn).real), float(np.vdot(r, r).real)))
        x = xn
        if stop and np.sqrt(stats[-1][0] / stats[-1][1]) < relative_variation and np.sqrt(stats[-1][2]) < res_tol:
            break
    return x, np.array(stats), gamma, it


def forward_backward(op: "Operator", entries, y, iters, sigma, mu=None, beta=None, tight_frame=True,
                     l1_itermax=50, l1_tol=1e-2, l1_nu=1.0, positivity=True, real=True):
    """Forward-backward splitting (Combettes & Wajs 2005) for min ||y - Phi x||^2 / (2 sigma^2) +
    mu ||Psi^dagger x||_1 [+ i_C], in the PhiTPhi form cpp/example/fb_random_coverage.cc:89-106 sets up
    (target the dirty image Phi^dagger y), FP64:
        x~ = x - (beta / sigma^2)(Phi^dagger Phi x - Phi^dagger y);  x = prox_(mu beta ||Psi^dagger .||_1)(x~)
    from x = 0; the proximal is the tight-frame closed form or l1_prox_constrained.  mu / beta default to
    the example's 1e-5 N and 0.5 / N.  SOPT (not in the reference tree) is not followed line by line:
    parity with its iterates is unpinned.  Returns (x, stats): stats[k] = (||x_k - x_(k-1)||^2,
    ||x_k||^2, ||Phi^dagger Phi x_(k-1) - Phi^dagger y||^2)."""
    ny, nx = op.ny, op.nx
    N = ny * nx
    mu = 1e-5 * N if mu is None or mu <= 0 else mu
    beta = 0.5 / N if beta is None or beta <= 0 else beta
    step, gam = beta / sigma ** 2, mu * beta
    psit = lambda v: sara_analysis(v, ny, nx, entries)  # noqa: E731
    psi = lambda c: sara_synthesis(c, ny, nx, entries)  # noqa: E731
    dimage = op.adjoint(np.asarray(y, np.complex128))
    x = np.zeros(N, np.complex128)
    stats = []
    for _ in range(iters):
        res = op.adjoint(op.forward(x)) - dimage
        xu = x - step * res
        if tight_frame:
            c = psit(xu)
            xn = xu + psi(soft_threshold(c, gam * l1_nu) - c) / l1_nu
        else:
            xn, _ = l1_prox_constrained(xu, gam, psit, psi, l1_nu, l1_itermax, l1_tol, positivity, real)
        stats.append((float(np.vdot(xn - x, xn - x).real), float(np.vdot(xn, xn).real), float(np.vdot(res, res).real)))
        x = xn
    return x, np.array(stats)


def snr_db(model, solution):
    """20 log10(||model|| / ||model - solution||)."""
    model, solution = np.asarray(model), np.asarray(solution)
    return float(20 * np.log10(np.linalg.norm(model) / np.linalg.norm(model - solution)))


def omp_threads():
    return lib().orc_omp_threads()


def set_omp_threads(n):
    lib().orc_set_omp_threads(int(n))
