"""wproj_utilities (cpp/purify/wproj_utilities.h / .cc): the chirp-based w-projection matrix.

wprojection_matrix(G, x_size, y_size, w_components, cell_x, cell_y, energy_fraction_chirp,
energy_fraction_wproj) mirrors wproj_utilities::wprojection_matrix (wproj_utilities.cc:132-188):
every row of the sparse gridding matrix G is convolved with the hard-thresholded FFT of its w chirp
and hard-thresholded again.  The work runs on the GPU in FP64; G and the result are scipy CSR
matrices of complex128.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib


def to_pf_csr(G):
    """A pf_csr view of a scipy sparse matrix (complex128, sorted indices) and the arrays it points to."""
    G = sp.csr_matrix(G, dtype=np.complex128)
    G.sort_indices()
    rowptr = np.ascontiguousarray(G.indptr, np.int64)
    col = np.ascontiguousarray(G.indices, np.int32)
    vals = np.ascontiguousarray(G.data, np.complex128)
    m = _lib.pf_csr()
    m.rows, m.cols = G.shape
    m.nnz = G.nnz
    m.rowptr = rowptr.ctypes.data_as(C.POINTER(C.c_int64))
    m.colidx = col.ctypes.data_as(C.POINTER(C.c_int32))
    m.vals = vals.ctypes.data_as(C.POINTER(C.c_double))
    return m, (rowptr, col, vals)


def wprojection_matrix(G, x_size, y_size, w_components, cell_x, cell_y, energy_fraction_chirp=1.0,
                       energy_fraction_wproj=1.0):
    m, keep = to_pf_csr(G)
    w = np.ascontiguousarray(w_components, np.float64).ravel()
    if w.size != m.rows:
        raise RuntimeError("w_components must have one entry per row of G")
    out = _lib.pf_csr()
    _lib.check(_lib.lib().pf_wprojection_matrix(C.byref(m), int(x_size), int(y_size), w.ctypes.data, float(cell_x),
                                                float(cell_y), float(energy_fraction_chirp),
                                                float(energy_fraction_wproj), C.byref(out)))
    try:
        rows, cols, nnz = out.rows, out.cols, out.nnz
        rp = np.ctypeslib.as_array(out.rowptr, (rows + 1,)).copy()
        ci = np.ctypeslib.as_array(out.colidx, (max(nnz, 1),))[:nnz].copy()
        vv = np.ctypeslib.as_array(out.vals, (2 * max(nnz, 1),))[:2 * nnz].copy().view(np.complex128)
    finally:
        _lib.lib().pf_csr_free(C.byref(out))
    del keep
    return sp.csr_matrix((vv, ci, rp), shape=(rows, cols))
