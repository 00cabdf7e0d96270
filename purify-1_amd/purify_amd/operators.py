"""Python mirror of PURIFY's measurement-operator interface on top of the HIP C-ABI.

    op = measurementoperator.init_degrid_operator_2d(u, v, w, weights, imsizey, imsizex,
                                                     oversample_ratio=2, power_iters=100, power_tol=1e-4,
                                                     kernel=kernels.kernel.kb, Ju=4, Jv=4, w_term=False,
                                                     Jw=6, cellx=1, celly=1)
    y = op * x                 # Phi x
    x = op.adjoint() * y       # Phi^dagger y

mirrors measurementoperator::init_degrid_operator_2d (cpp/purify/operators.h:373-414) and its
sopt::LinearTransform result.  Vectors may be numpy arrays (host, complex128, as the reference's
Vector<t_complex>) or torch CUDA tensors in the operator's precision (complex64 / complex128), in
which case the apply runs on the tensor's memory on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .kernels import kernel as _kernel
from .utilities import vis_params, vis_units

PRECISIONS = {"c64": _lib.PF_C64, "c128": _lib.PF_C128, "complex64": _lib.PF_C64, "complex128": _lib.PF_C128}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _torch():
    import torch
    return torch


def _is_device_tensor(x):
    try:
        torch = _torch()
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


def check_device_tensor(t, n, dtype, device=None, what="tensor"):
    """Raise unless t is a contiguous CUDA tensor of n entries of dtype (on device, if given).

    The C library reads and writes device memory through the raw pointer with the operator's sizes, so
    every device tensor that crosses the boundary is checked here first."""
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError(f"{what} must be a CUDA tensor")
    if t.dtype != dtype:
        raise TypeError(f"{what} must be {dtype}, got {t.dtype}")
    if device is not None and t.device != device:
        raise ValueError(f"{what} is on {t.device}, expected {device}")
    if t.numel() != n:
        raise ValueError(f"{what} has {t.numel()} entries, expected {n}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    return t


class MeasurementOperator:
    """Owns a pf_op handle; the direct (Phi) and adjoint (Phi^dagger) applies of one operator."""

    def __init__(self, desc: _lib.pf_op_desc, keep=(), csr=None):
        self._keep = keep
        self._streams = _lib.StreamLog()
        h = C.c_void_p()
        if csr is None:
            _lib.check(_lib.lib().pf_op_create(C.byref(desc), C.byref(h)))
        else:
            _lib.check(_lib.lib().pf_op_create_csr(C.byref(desc), C.byref(csr), C.byref(h)))
        self._h = h
        info = self.info()
        self.imsizey, self.imsizex = info.imsizey, info.imsizex
        self.M = info.M
        self.precision = info.precision
        self.grid_shape = (info.Nv, info.Nu)

    # -- lifetime
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._streams.sync()
            _lib.lib().pf_op_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- info / norm
    def info(self):
        inf = _lib.pf_op_info()
        _lib.check(_lib.lib().pf_op_get_info(self._h, C.byref(inf)))
        return inf

    @property
    def norm(self):
        d = C.c_double()
        _lib.check(_lib.lib().pf_op_get_norm(self._h, C.byref(d)))
        return d.value

    def set_norm(self, norm):
        _lib.check(_lib.lib().pf_op_set_norm(self._h, float(norm)))

    def power_method(self, iters, tol, start=None):
        """||Phi||_2 of the operator as it is now (SOPT power_method, operators.h:390-391)."""
        out = C.c_double()
        if start is None:
            _lib.check(_lib.lib().pf_op_power_method(self._h, int(iters), float(tol), None, C.byref(out)))
        else:
            s = _c128(start)
            _lib.check(_lib.lib().pf_op_power_method(self._h, int(iters), float(tol), s.ctypes.data, C.byref(out)))
        return out.value

    # -- tuning
    def set_tuning(self, key, value):
        _lib.check(_lib.lib().pf_op_set_tuning(self._h, key.encode(), int(value)))

    def get_tuning(self, key):
        v = C.c_int64()
        _lib.check(_lib.lib().pf_op_get_tuning(self._h, key.encode(), C.byref(v)))
        return v.value

    def autotune(self, reps=3):
        """Time the gridding-kernel variants on this operator and keep the fastest; returns it."""
        v = C.c_int64()
        _lib.check(_lib.lib().pf_op_autotune(self._h, int(reps), C.byref(v)))
        return v.value

    # -- per-kernel timing
    def profile(self, on=True):
        """Enable (and clear) or disable HIP-event timing of every kernel this operator launches."""
        _lib.check(_lib.lib().pf_op_profile_enable(self._h, int(bool(on))))

    def profile_table(self):
        """{kernel name: (total ms, launches)} since profiling was enabled."""
        out = {}
        name = C.create_string_buffer(64)
        tot = C.c_double()
        cnt = C.c_int64()
        i = 0
        while _lib.lib().pf_op_profile_get(self._h, i, name, 64, C.byref(tot), C.byref(cnt)) == _lib.PF_OK:
            out[name.value.decode()] = (tot.value, cnt.value)
            i += 1
        return out

    # -- applies
    def _dtype(self):
        torch = _torch()
        return torch.complex64 if self.precision == _lib.PF_C64 else torch.complex128

    def _apply(self, adjoint, x, out=None, stream=None):
        nin = self.M if adjoint else self.imsizex * self.imsizey
        nout = self.imsizex * self.imsizey if adjoint else self.M
        fn = _lib.lib().pf_op_adjoint if adjoint else _lib.lib().pf_op_forward
        if _is_device_tensor(x):
            torch = _torch()
            if x.dtype != self._dtype():
                raise TypeError(f"device input must be {self._dtype()} for this operator")
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
        xh = _c128(x).ravel()
        if xh.size != nin:
            raise ValueError(f"input has {xh.size} entries, expected {nin}")
        yh = np.empty(nout, np.complex128)
        _lib.check(fn(self._h, xh.ctypes.data, yh.ctypes.data, _lib.PF_HOST, _lib.PF_HOST, None))
        return yh

    def _call(self, fn, handle, x, nin, nout, out=None, stream=None):
        if _is_device_tensor(x):
            torch = _torch()
            if x.dtype != self._dtype():
                raise TypeError(f"device input must be {self._dtype()} for this operator")
            x = x.contiguous()
            if x.numel() != nin:
                raise ValueError(f"input has {x.numel()} entries, expected {nin}")
            if out is None:
                out = torch.empty(nout, dtype=x.dtype, device=x.device)
            else:
                check_device_tensor(out, nout, x.dtype, x.device, "output")
            s = stream if stream is not None else torch.cuda.current_stream(x.device).cuda_stream
            self._streams.add(s, x.device)
            _lib.check(fn(handle, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), _lib.PF_DEVICE,
                          _lib.PF_DEVICE, C.c_void_p(s)))
            return out
        xh = _c128(x).ravel()
        if xh.size != nin:
            raise ValueError(f"input has {xh.size} entries, expected {nin}")
        if out is not None:  # sopt::OperatorFunction's output-first form: fill the caller's array
            if not (isinstance(out, np.ndarray) and out.dtype == np.complex128 and out.size == nout
                    and out.flags.c_contiguous):
                raise TypeError(f"output must be a contiguous complex128 array of {nout} entries")
            yh = out.reshape(-1)
        else:
            yh = np.empty(nout, np.complex128)
        _lib.check(fn(handle, xh.ctypes.data, yh.ctypes.data, _lib.PF_HOST, _lib.PF_HOST, None))
        return out if out is not None else yh

    def normal(self, x, out=None, stream=None):
        """Phi^dagger Phi x (the compact grid-degrid operator; see init_grid_degrid_operator_2d)."""
        n = self.imsizex * self.imsizey
        return self._call(_lib.lib().pf_op_apply_normal, self._h, x, n, n, out, stream)

    def forward(self, x, out=None, stream=None):
        return self._apply(False, x, out, stream)

    def adjoint_apply(self, y, out=None, stream=None):
        return self._apply(True, y, out, stream)

    # -- stage access (device tensors, op precision; no norm)
    def _stage(self, name, x, nin, nout):
        torch = _torch()
        if not _is_device_tensor(x):
            raise TypeError("stage access takes a CUDA tensor")
        x = check_device_tensor(x.contiguous(), nin, self._dtype(), None, "input")
        out = torch.empty(nout, dtype=x.dtype, device=x.device)
        s = torch.cuda.current_stream(x.device).cuda_stream
        self._streams.add(s, x.device)
        _lib.check(getattr(_lib.lib(), name)(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()),
                                             C.c_void_p(s)))
        return out

    def apply_G(self, grid):
        ng = self.grid_shape[0] * self.grid_shape[1]
        return self._stage("pf_op_apply_G", grid, ng, self.M)

    def apply_GH(self, y):
        return self._stage("pf_op_apply_GH", y, self.M, self.grid_shape[0] * self.grid_shape[1])

    def apply_FZ(self, x):
        return self._stage("pf_op_apply_FZ", x, self.imsizex * self.imsizey, self.grid_shape[0] * self.grid_shape[1])

    def apply_FZ_adjoint(self, grid):
        ng = self.grid_shape[0] * self.grid_shape[1]
        return self._stage("pf_op_apply_FZ_adjoint", grid, ng, self.imsizex * self.imsizey)


class _Adjoint:
    def __init__(self, lt):
        self._lt = lt

    def __mul__(self, y):
        return self._lt.op.adjoint_apply(y)

    __call__ = __mul__

    def adjoint(self):
        return self._lt

    @property
    def sizes(self):
        return self._lt.sizes[::-1]


class LinearTransform:
    """sopt::LinearTransform-like pair: `lt * x` is Phi x, `lt.adjoint() * y` is Phi^dagger y."""

    def __init__(self, op: MeasurementOperator):
        self.op = op
        self.sizes = (op.M, op.imsizex * op.imsizey)  # (rows, cols)

    def __mul__(self, x):
        return self.op.forward(x)

    __call__ = __mul__

    def adjoint(self):
        return _Adjoint(self)

    @property
    def norm(self):
        return self.op.norm


def _make_desc(u, v, w, weights, imsizey, imsizex, oversample_ratio, power_iters, power_tol, kernel, Ju, Jv,
               w_term, Jw, cellx, celly, units, precision, tile=(0, 0), wproj=None):
    u, v = _f64(u).ravel(), _f64(v).ravel()
    M = u.size
    if v.size != M:
        raise RuntimeError("Size of u and v vectors are not the same for creating gridding matrix.")
    w = np.zeros(M) if w is None else _f64(w).ravel()
    if w_term and w.size != M:
        raise RuntimeError("Size of u and w vectors are not the same for creating gridding matrix.")
    wt = None if weights is None else _c128(weights).ravel()
    if wt is not None and wt.size != M:
        raise RuntimeError("Size of u and w vectors are not the same for creating gridding matrix.")
    d = _lib.pf_op_desc()
    _lib.lib().pf_op_desc_init(C.byref(d))
    d.imsizey, d.imsizex = int(imsizey), int(imsizex)
    d.oversample_ratio = float(oversample_ratio)
    d.kernel = int(_kernel.coerce(kernel))
    d.Ju, d.Jv = int(Ju), int(Jv)
    d.w_term, d.Jw = int(bool(w_term)), int(Jw)
    d.cellx, d.celly = float(cellx), float(celly)
    d.units = int(units)
    d.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    d.M = M
    d.u, d.v, d.w = u.ctypes.data, v.ctypes.data, w.ctypes.data
    d.weights = None if wt is None else wt.ctypes.data
    d.power_iters = int(power_iters)
    d.power_tol = float(power_tol)
    d.tile_u, d.tile_v = int(tile[0]), int(tile[1])
    if w_term and wproj is not None and wproj.get("w_projection", "approx") == "chirp":
        d.w_term = 2
        d.energy_chirp = float(wproj.get("energy_chirp", 0.999))
        d.energy_wproj = float(wproj.get("energy_wproj", 0.999))
    elif wproj is not None and wproj.get("w_projection", "approx") not in ("approx", "chirp"):
        raise ValueError("w_projection must be 'approx' or 'chirp'")
    return d, (u, v, w, wt)


class measurementoperator:  # noqa: N801 -- mirrors the reference's namespace
    @staticmethod
    def init_degrid_operator_2d(*args, **kw):
        """Both reference overloads:
        init_degrid_operator_2d(u, v, w, weights, imsizey, imsizex, oversample_ratio=2, power_iters=100,
                                power_tol=1e-4, kernel=kb, Ju=4, Jv=4, w_term=False, Jw=6, cellx=1, celly=1)
            -- u, v in pixels (operators.h:373-396)
        init_degrid_operator_2d(uv_vis, imsizey, imsizex, cell_x, cell_y, oversample_ratio=2, power_iters=100,
                                power_tol=1e-4, kernel=kb, Ju=4, Jv=4, w_term=False, Jw=6)
            -- vis_params in lambda / radians / pixels (operators.h:398-414)
        Extra keywords: precision ("c64" default, or "c128"), tile=(tile_u, tile_v) or tile="auto" (the shape,
        and for complex64 the permutation bucket, chosen by timing Phi + Phi^dagger pairs on one build per
        candidate; read back with op.get_tuning("tile_u" / "tile_v" / "perm_bucket")); with w_term,
        w_projection="approx" (the projection kernels of operators.cc:49-99, the default) or "chirp" (the rows of
        wproj_utilities::wprojection_matrix applied to the plain gridding matrix, wproj_utilities.cc:132-188,
        built on the device; square grids) with energy_chirp and energy_wproj (0.999 each by default).
        """
        precision = kw.pop("precision", "c64")
        tile = kw.pop("tile", (0, 0))
        if isinstance(tile, str):
            if tile != "auto":
                raise ValueError('tile must be (tile_u, tile_v) or "auto"')
            tile = (-1, -1)
        wproj = {k: kw.pop(k) for k in ("w_projection", "energy_chirp", "energy_wproj") if k in kw}
        if args and isinstance(args[0], vis_params):
            return _init_from_vis(*args, precision=precision, tile=tile, wproj=wproj, **kw)
        return _init_raw(*args, precision=precision, tile=tile, wproj=wproj, **kw)


def _init_raw(u, v, w, weights, imsizey, imsizex, oversample_ratio=2.0, power_iters=100, power_tol=1e-4,
              kernel=_kernel.kb, Ju=4, Jv=4, w_term=False, Jw=6, cellx=1.0, celly=1.0, precision="c64", tile=(0, 0),
              wproj=None):
    d, keep = _make_desc(u, v, w, weights, imsizey, imsizex, oversample_ratio, power_iters, power_tol, kernel, Ju,
                         Jv, w_term, Jw, cellx, celly, _lib.PF_PIXELS, precision, tile, wproj)
    return LinearTransform(MeasurementOperator(d, keep))


def _init_from_vis(uv_vis, imsizey, imsizex, cell_x, cell_y, oversample_ratio=2.0, power_iters=100,
                   power_tol=1e-4, kernel=_kernel.kb, Ju=4, Jv=4, w_term=False, Jw=6, precision="c64", tile=(0, 0),
                   wproj=None):
    units = {vis_units.lambda_: _lib.PF_LAMBDA, vis_units.radians: _lib.PF_RADIANS,
             vis_units.pixels: _lib.PF_PIXELS}[uv_vis.units]
    d, keep = _make_desc(uv_vis.u, uv_vis.v, uv_vis.w, uv_vis.weights, imsizey, imsizex, oversample_ratio,
                         power_iters, power_tol, kernel, Ju, Jv, w_term, Jw, cell_x, cell_y, units, precision, tile,
                         wproj)
    return LinearTransform(MeasurementOperator(d, keep))


def init_degrid_operator_2d_from_matrix(G, imsizey, imsizex, oversample_ratio=2.0, power_iters=100, power_tol=1e-4,
                                        kernel=_kernel.kb, Ju=4, Jv=4, precision="c64"):
    """Phi = G F Z S / norm with a given sparse gridding matrix G (rows = visibilities, columns = grid
    cells p * Nu + q), e.g. wproj_utilities.wprojection_matrix's result.  The kernel and supports fix
    the correction S (operators.cc:101-121); power_iters / power_tol the normalisation
    (operators.h:390-391)."""
    from .wproj_utilities import to_pf_csr
    csr, keep = to_pf_csr(G)
    z = np.zeros(0)
    d, keep2 = _make_desc(z, z, None, None, imsizey, imsizex, oversample_ratio, power_iters, power_tol, kernel, Ju,
                          Jv, False, 6, 1.0, 1.0, _lib.PF_PIXELS, precision)
    return LinearTransform(MeasurementOperator(d, (keep, keep2), csr=csr))


measurementoperator.init_degrid_operator_2d_from_matrix = staticmethod(init_degrid_operator_2d_from_matrix)


class SuperOperator:
    """measurementoperator::init_combine_operators (operators.h:513-527) on the device: ``sup * x`` stacks
    Phi_b x over the blocks, ``sup.adjoint() * y`` is sum_b Phi_b^dagger y_b.  The blocks share one F Z S
    per apply (pf_super_*).  Inputs: torch CUDA tensors in the operators' precision, or host arrays
    (copied through the device)."""

    def __init__(self, transforms):
        self.transforms = list(transforms)  # kept alive: the handle uses their operators
        ops = [t.op if isinstance(t, LinearTransform) else t for t in self.transforms]
        arr = (C.c_void_p * len(ops))(*[o.handle.value if isinstance(o.handle, C.c_void_p) else o.handle for o in ops])
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pf_super_create(arr, len(ops), C.byref(self._h)))
        m = C.c_int64()
        _lib.check(_lib.lib().pf_super_size(self._h, C.byref(m)))
        self.M = int(m.value)
        self.segments = []
        o = 0
        for op in ops:
            self.segments.append((o, op.M))
            o += op.M
        self.ops = ops
        self.imsizey, self.imsizex = ops[0].imsizey, ops[0].imsizex
        self.sizes = (self.M, self.imsizex * self.imsizey)
        self._streams = _lib.StreamLog()

    def close(self):
        if self._h:
            self._streams.sync()
            _lib.lib().pf_super_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _run(self, adjoint, x):
        torch = _torch()
        dt = self.ops[0]._dtype()
        nin = self.M if adjoint else self.imsizex * self.imsizey
        nout = self.imsizex * self.imsizey if adjoint else self.M
        host = not _is_device_tensor(x)
        if host:
            xh = _c128(x).ravel()
            xd = torch.from_numpy(xh).to(dt).cuda()
        else:
            if x.dtype != dt:
                raise TypeError(f"device input must be {dt} for these operators")
            xd = x.contiguous()
        if xd.numel() != nin:
            raise ValueError(f"input has {xd.numel()} entries, expected {nin}")
        out = torch.empty(nout, dtype=dt, device=xd.device)
        s = torch.cuda.current_stream(xd.device).cuda_stream
        self._streams.add(s, xd.device)
        fn = _lib.lib().pf_super_adjoint if adjoint else _lib.lib().pf_super_forward
        _lib.check(fn(self._h, C.c_void_p(xd.data_ptr()), C.c_void_p(out.data_ptr()), C.c_void_p(s)))
        if host:
            return out.to(torch.complex128).cpu().numpy()
        return out

    def __mul__(self, x):
        return self._run(False, x)

    __call__ = __mul__

    def adjoint(self):
        sup = self

        class _A:
            def __mul__(self, y):
                return sup._run(True, y)

            __call__ = __mul__
        return _A()


def init_combine_operators(measure_ops, seg=None):
    """operators.h:513-527.  ``seg`` (the stacked vector's (start, size) per block) is implied by the
    operators' sizes in block order; given, it must agree."""
    sup = SuperOperator(measure_ops)
    if seg is not None and [tuple(map(int, s)) for s in seg] != sup.segments:
        raise RuntimeError("Numer of groups and measurement operators does not match.")
    return sup


def _segments(total, seg_num):
    """operators.h:562-576: seg_num blocks of floor(total / seg_num), the last taking the remainder."""
    if seg_num == 0:
        raise RuntimeError("Number of segments = 0")
    size = int(np.floor(total / seg_num))
    seg, t = [], 0
    for i in range(seg_num):
        d = total - t if i == seg_num - 1 else size
        seg.append((t, d))
        t += d
    if t != total:
        raise RuntimeError("Segments do not sum to the total")
    return seg


def _vis_segment(uv, start, size):
    from .utilities import vis_params
    sl = slice(start, start + size)
    return vis_params(u=uv.u[sl], v=uv.v[sl], w=uv.w[sl], vis=uv.vis[sl], weights=uv.weights[sl], units=uv.units,
                      ra=uv.ra, dec=uv.dec, average_frequency=uv.average_frequency)


def init_super_operator(create_operator, data, *args, **kw):
    """operators.h:531-578, three overloads: (create_operator, [vis_params, ...], args...),
    (create_operator, [(start, size), ...], vis_params, args...) and (create_operator, vis_params, seg_num,
    args...).  create_operator(uv_block, *args, **kw) builds each block's operator (normalised by its own
    power method, as in the reference)."""
    from .utilities import vis_params
    if isinstance(data, vis_params):
        seg_num, args = int(args[0]), args[1:]
        blocks = [_vis_segment(data, a, b) for a, b in _segments(data.size(), seg_num)]
    elif len(data) and isinstance(data[0], vis_params):
        blocks = list(data)
    else:
        uv, args = args[0], args[1:]
        blocks = [_vis_segment(uv, int(a), int(b)) for a, b in data]
    ops = [create_operator(b, *args, **kw) for b in blocks]
    if len(ops) != len(blocks):
        raise RuntimeError("Numer of groups and measurement operators does not match.")
    return init_combine_operators(ops)


measurementoperator.init_combine_operators = staticmethod(init_combine_operators)
measurementoperator.init_super_operator = staticmethod(init_super_operator)


class GridDegridOperator:
    """operators::init_grid_degrid_operator_2d (compact_operators.h:150-171): the self-adjoint operator
    x -> Phi^dagger Phi x, normalised by the square of the measurement operator's norm.  The reference
    returns a sopt::OperatorFunction (output first, input second); here ``op(x)`` returns the output and
    ``op(out, x)`` fills ``out``.  Wrap it as ``LinearTransform``-like pair with the identity adjoint
    exactly as the reference's tests do."""

    def __init__(self, mop: MeasurementOperator):
        self.op = mop

    def __call__(self, *args, stream=None):
        if len(args) == 2:
            out, x = args
            return self.op.normal(x, out=out, stream=stream)
        (x,) = args
        return self.op.normal(x, stream=stream)

    @property
    def norm(self):
        return self.op.norm


class PsfOperator:
    """operators::init_psf_convolve_2d (compact_operators.h:279-306): x -> IFFT(FFT(psf) . FFT(x)) / norm^2
    with psf = Phi^dagger Phi delta and norm from its own power method."""

    def __init__(self, mop: MeasurementOperator, power_iters=100, power_tol=1e-4, start=None):
        self.op = mop  # keeps the measurement operator alive
        self._h = C.c_void_p()
        st = None if start is None else _c128(start).ravel()
        if st is not None and st.size != mop.imsizex * mop.imsizey:
            raise ValueError("start vector has the wrong size")
        self._keep = st
        _lib.check(_lib.lib().pf_psf_create(mop.handle, int(power_iters), float(power_tol),
                                            None if st is None else st.ctypes.data, C.byref(self._h)))

    def close(self):
        if self._h:
            self.op._streams.sync()  # the PSF's applies log their streams with the measurement operator
            _lib.lib().pf_psf_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def norm(self):
        v = C.c_double()
        _lib.check(_lib.lib().pf_psf_norm(self._h, C.byref(v)))
        return v.value

    def __call__(self, *args, stream=None):
        n = self.op.imsizex * self.op.imsizey
        if len(args) == 2:
            out, x = args
            return self.op._call(_lib.lib().pf_psf_apply, self._h, x, n, n, out, stream)
        (x,) = args
        return self.op._call(_lib.lib().pf_psf_apply, self._h, x, n, n, None, stream)


class operators:  # noqa: N801 -- mirrors the reference's purify::operators namespace (compact operators)
    @staticmethod
    def init_grid_degrid_operator_2d(*args, **kw):
        """Both reference overloads (compact_operators.h:150-189):
        init_grid_degrid_operator_2d(u, v, w, weights, imsizey, imsizex, oversample_ratio=2, power_iters=100,
                                     power_tol=1e-4, kernel=kb, Ju=4, Jv=4, w_term=False, Jw=6, cellx=1, celly=1)
        init_grid_degrid_operator_2d(uv_vis, imsizey, imsizex, cell_x, cell_y, oversample_ratio=2, power_iters=100,
                                     power_tol=1e-4, kernel=kb, Ju=4, Jv=4, w_term=False, Jw=6)
        Extra keyword: precision ("c64" default, or "c128")."""
        lt = measurementoperator.init_degrid_operator_2d(*args, **kw)
        return GridDegridOperator(lt.op)

    @staticmethod
    def init_psf_convolve_2d(measure_op, imsizey=None, imsizex=None, power_iters=100, power_tol=1e-4, start=None):
        """init_psf_convolve_2d(degrid_grid, imsizey, imsizex, power_iters, power_tol)
        (compact_operators.h:279-306), from a measurement operator's LinearTransform."""
        mop = measure_op.op if isinstance(measure_op, LinearTransform) else measure_op
        if imsizey is not None and (imsizey, imsizex) != (mop.imsizey, mop.imsizex):
            raise ValueError("image size differs from the measurement operator's")
        return PsfOperator(mop, power_iters, power_tol, start)


# factory::measurement_operator_factory (cpp/purify/measurement_operator_factory.h:20-96)
class distributed_measurement_operator:  # noqa: N801
    serial = "serial"
    gpu_serial = "gpu_serial"
    mpi_distribute_image = "mpi_distribute_image"
    mpi_distribute_grid = "mpi_distribute_grid"
    gpu_mpi_distribute_image = "gpu_mpi_distribute_image"
    gpu_mpi_distribute_grid = "gpu_mpi_distribute_grid"


def measurement_operator_factory(dist, *args, **kw):
    """Builds the operator on the GPU for `serial` and `gpu_serial`; the MPI variants have no
    single-GPU meaning here and raise, as the reference does when built without MPI."""
    if dist in (distributed_measurement_operator.serial, distributed_measurement_operator.gpu_serial):
        return measurementoperator.init_degrid_operator_2d(*args, **kw)
    raise RuntimeError("Compile with MPI if you want to use MPI distributed operators.")
